// basis.hip -- symmetry functions G (NNP::compute_symm, src/nnp/nnp.cpp:350-392), their adjoint dE/dd
// for every neighbour pair (the geometric half of NNP::compute_force, src/nnp/nnp.cpp:412-476), and
// their tangent along per-atom displacements (dG/dr recomputed for the force-matching gradient).
//
// One workgroup (BW = 4 or 2 wavefronts) per centre atom; the real type T (float or double) is a template
// parameter: on gfx950 without packed-FP32 code the vector ALU runs FP64 FMAs at the FP32 rate, and the
// triplet loop below has no transcendental, so T = double costs registers and LDS, not ALU rate.
//
// Wave 0 buckets the atom's neighbours inside the basis cutoff by species (ballot order = list order,
// so the slot assignment is deterministic); all threads then fill LDS with each slot's unit vector and
// distance, its radial cutoff, and, per distinct angular eta, the separable factor
// E_s = fc(r_s) exp(-r_s^2/eta^2) with its r-derivative, so a GAUSS triplet term
//   fc(rij) fc(rik) |1/2 + lambda/2 cos|^zeta exp(-(rij^2+rik^2)/eta^2)      (basis_angular.cpp:371-382)
// costs one product E_j E_k.  Window powers come from a per-triplet cache of (1 +- cos)/2 raised to
// 2^k and 2^k - 1, so a power-of-two zeta is one register pick.  Triplets of an angular block (J,K) --
// the rectangle J x K, or the triangle j<k of J x J -- are spread over the workgroup's threads; functions are
// kept in registers and reduced wave by wave, the wave partials added in a fixed order.  Radial
// functions map one function to one lane and neighbour groups to lane groups.  The adjoint accumulates
// dE/dd per slot in one LDS array per wave and sums them in a fixed order: results are
// bit-reproducible.
#include "basis_math.hpp"

// BW (a template parameter of the kernel and its helpers) is the number of wavefronts per workgroup, i.e.
// per atom, and BT = 64 BW its thread count: 4 by default; 2 halves the per-wave accumulators of the
// gradient modes, which lets more atoms share a CU where LDS limits residency.

// Modes: 0 = network input rows inpw*(G + inpb); 1 = adjoint (per-pair dE/dd from dE/dG rows);
// 2 = tangent rows v = sum_a eps_a . dG/dr_a; 3 = fused (0, the network, then 1).
// Dynamic LDS per workgroup (nbm = neighbour slots, kp = kmax rounded up to 4, nrp = the most radial
// functions of a species):
//   g[BW 3 nbm] (FP64, modes 1 and 3: per-wave dE/dd accumulators; LDS FP64 atomic adds run several
//         times faster than FP32 ones on gfx950);
//   FP64: row[kp] (G sums, then dE/dG; FP32 in the lean FP32 adjoint), red[nred], rp[2 nrp] (radial eta,
//         rs of the species), sc[2 kp] (mode 3: inpw, inpb);
//   T (16-byte aligned): u[4 nbm] (unit vector, 1/|d|), E[nbm][2 NG] (the angular factors E and their
//      r-derivatives Ed of a slot side by side; at least BW x 16 doubles, which the gradient epilogue
//      reuses for its partial sums), te[4 nbm] (mode 2), fR[nbm], gR[nbm], r[nbm], fr[nbm] (modes 1 and
//      3), fA[nbm], gA[nbm] (generic angular path: angular cutoff and derivative);
//   p[nbm] (int; 16-bit row positions in the lean layout), ci[kp] (bytes);  float: rlo[nbm] (T = float);
//   mode 3 appends the network area (netf floats: staged weights, activations, partial sums), 16-byte
//   aligned.
// The triplet loops read a slot's geometry and its angular factors as whole 16-byte records.
// red holds the per-wave partial sums (BW 64, FP64) and, in the adjoint, the dE/dG weights of every
// canonical angular block in (sign, level, group) order (14 ng FP32 values per TensorTri block).
// Each lane sums its share of a symmetry function in T, and so does the wave reduction; the sum over the
// waves is FP64, and so is the G row, which reaches the input scaling inpw (G + inpb) (also FP64) without the
// FP32 rounding that the mean removal would magnify by mean / deviation.
template <int BW>
__host__ __device__ inline int basis_nred(int nt, int ng, int mode) {
    const int wt = (nt * (nt + 1) / 2 * 14 * ng + 1) / 2;   // FP32 weight table, in doubles
    return mode == 1 ? wt : (mode == 3 ? max(BW * 64, wt) : BW * 64);
}

// byte offset of the T area: the FP64 arrays, rounded up to 16
template <int BW>
__host__ __device__ inline int basis_t_off(int nbm, int ng, int kmax, int nrp, int mode, int nt, bool rowf) {
    const int kp = (kmax + 3) & ~3;
    const bool gacc = (mode == 1 || mode == 3);
    const int gb = gacc ? 8 * BW * 3 * nbm : 0;
    const int nd = (rowf ? kp / 2 : kp) + basis_nred<BW>(nt, ng, mode) + 2 * nrp + (mode == 3 ? 2 * kp : 0);
    return (gb + 8 * nd + 15) & ~15;
}

// The split adjoint of an all-canonical model (mode 1 without the generic path) keeps no per-slot radial
// arrays: its per-slot pass adds each neighbour's radial dE/dd to wave 0's accumulators directly, and |d| in
// FP64 passes from the bucketing to that pass through wave 1's accumulators, which are cleared afterwards.
__host__ __device__ inline bool basis_lean(int mode, bool gen) { return mode == 1 && !gen; }

// The lean FP32 adjoint keeps its dE/dG row in FP32 (its source is FP32); the others in FP64.
__host__ __device__ inline bool basis_rowf(int mode, bool gen, int tsize) { return basis_lean(mode, gen) && tsize == 4; }

// Elements of the angular-factor table E: 2 ng per slot, and at least the epilogue's BW x 16 FP64 partial
// sums, which reuse it once the sums are done.
template <int BW>
__host__ __device__ inline int basis_ne(int nbm, int ng, int tsize) {
    const int n = 2 * ng * nbm, m = BW * 16 * 8 / tsize;
    return n > m ? n : m;
}

template <class T, int BW>
__host__ __device__ inline int basis_lds_bytes(int nbm, int ng, int kmax, int nrp, int mode, int netf, int nt, bool gen) {
    const int kp = (kmax + 3) & ~3;
    const bool gacc = (mode == 1 || mode == 3);
    const bool lean = basis_lean(mode, gen);
    int n = 4 * nbm + basis_ne<BW>(nbm, ng, (int)sizeof(T)) + (lean ? 0 : 3 * nbm);
    if (mode == 2) n += 4 * nbm;
    if (gacc && !lean) n += nbm;
    if (gen) n += 2 * nbm;
    const int nf = (sizeof(T) == 4 && !lean) ? nbm : 0;
    // slot records p (the lean layout keeps a 16-bit row position) and the canonical slot table ci (bytes)
    const int base = (basis_t_off<BW>(nbm, ng, kmax, nrp, mode, nt, basis_rowf(mode, gen, (int)sizeof(T))) +
                      (int)sizeof(T) * n + (lean ? 2 : 4) * nbm + kp + 4 * nf + 15) & ~15;
    return base + (mode == 3 ? 4 * netf : 0);
}

template <class T>
struct BasisLds {
    T* u;      // [nbm][4]: unit vector d / |d| and 1 / |d| (FP64 values rounded once)
    T* E;      // [nbm][2 NG]: fc exp(-r^2 / eta_g^2) of each slot and eta group, then its r derivative
    T* te;     // [nbm][4]
    T* fR;
    T* gR;
    T* r;      // |d|
    T* fr;     // modes 1 (not lean), 3: radial dE/dr of each slot
    T* fA;     // kernels with the generic angular path: angular cutoff of each slot and its r derivative
    T* gA;
    double* g; // [BW][3 nbm] per-wave dE/dd accumulators (one array per wave: a fixed order of adds)
    double* sred;  // [BW][16] the epilogue's partial sums (in E's area, free by then)
    float* rowf;   // the lean FP32 adjoint's dE/dG row
    double* row;
    double* red;   // [BW][64] partial sums; adjoint: canonical dE/dG weights [block][14 ng] as FP32
    double* rp;    // radial eta [nR], rs [nR] of the centre species
    double* sc;    // mode 3: network input scaling inpw [K], inpb [K]
    int* p;
    unsigned short* pe;   // lean layout: row position of the slot's pair
    unsigned char* ci;    // canonical slot of each angular function of the species
    float* rlo;    // T = float: |d| - (float)|d|, so the radial functions see |d| to FP64 precision
    float* nw; // mode 3: the species' network parameters (matrix-core layout), then activations
};

template <class T, int BW>
__device__ inline BasisLds<T> basis_lds(void* base, int nbm, int ng, int kmax, int nrp, int mode, int nt, bool gen) {
    BasisLds<T> L;
    const int kp = (kmax + 3) & ~3;
    const bool gacc = (mode == 1 || mode == 3);
    const bool rowf = basis_rowf(mode, gen, (int)sizeof(T));
    L.g = (double*)base;
    double* d = (double*)((char*)base + (gacc ? 8 * BW * 3 * nbm : 0));
    L.row = d;
    L.rowf = (float*)d;
    d += rowf ? kp / 2 : kp;
    L.red = d; d += basis_nred<BW>(nt, ng, mode);
    L.rp = d; d += 2 * nrp;
    L.sc = d;
    T* c = (T*)((char*)base + basis_t_off<BW>(nbm, ng, kmax, nrp, mode, nt, rowf));
    L.u = c; c += 4 * nbm;
    L.E = c;
    L.sred = (double*)c;
    c += basis_ne<BW>(nbm, ng, (int)sizeof(T));
    L.te = c; if (mode == 2) c += 4 * nbm;
    const bool lean = basis_lean(mode, gen);
    L.fR = c; if (!lean) c += nbm;
    L.gR = c; if (!lean) c += nbm;
    L.r = c; if (!lean) c += nbm;
    L.fr = c; if (gacc && !lean) c += nbm;
    L.fA = c; if (gen) c += nbm;
    L.gA = c; if (gen) c += nbm;
    L.p = (int*)c;
    L.pe = (unsigned short*)c;
    L.ci = (unsigned char*)c + (lean ? 2 : 4) * nbm;
    L.rlo = (float*)(L.ci + kp);
    float* e = L.rlo + ((sizeof(T) == 4 && !lean) ? nbm : 0);
    const long long end = ((long long)((char*)e - (char*)base) + 15) & ~15ll;
    L.nw = (float*)((char*)base + end);
    return L;
}

// |d| of slot q to FP64 precision
template <class T>
__device__ inline double slot_r64(const BasisLds<T>& L, int q) {
    if constexpr (sizeof(T) == 8) return (double)L.r[q];
    else return (double)L.r[q] + (double)L.rlo[q];
}

// (a, c) of flattened triplet t in block (J,K): rectangle a in J, c in K; triangle a<c in J.
// Triplet t of a block -> neighbour slots (a, c) by rotation: row d pairs x with (x + d) mod n, so the
// lanes of one row touch distinct a and distinct c (no LDS address shared by two lanes of an atomic).
//   same species (n = nJ):  d = 1 + t / n, a = t mod n, c = (a + d) mod n; rows d <= (n - 1) / 2 are full
//                           and, for even n, row n / 2 holds a < n / 2 only -- n (n - 1) / 2 triplets.
//   J != K (m = min, M = max of nJ, nK): d = t / m, x = t mod m, y = (x + d) mod M -- nJ nK triplets.
// rn = 1 / n (same) or 1 / m.
__device__ inline void tri_decode(int t, bool same, int nJ, int nK, float rn, int& a, int& c) {
    const int m = same ? nJ : min(nJ, nK), M = same ? nJ : max(nJ, nK);
    int d = (int)(((float)t + 0.5f) * rn);
    const int x0 = t - d * m;
    d += (x0 >= m) - (x0 < 0);
    const int x = t - d * m;
    int y = x + d + (same ? 1 : 0);
    y -= (y >= M) ? M : 0;
    const bool sw = !same && nJ > nK;
    a = sw ? y : x;
    c = sw ? x : y;
}

// pick v[g] for a workgroup-uniform g < ACE_MAXG
template <class T>
__device__ inline T pick4(const T* v, int g) {
    T r = v[0];
    r = (g == 1) ? v[1] : r;
    r = (g == 2) ? v[2] : r;
    r = (g == 3) ? v[3] : r;
    return r;
}

// c[s] for a workgroup-uniform s < 14 (a scalar branch, no indexed register access)
template <class T>
__device__ inline T sel14(const T* c, int s) {
    switch (s) {
        case 0: return c[0];   case 1: return c[1];   case 2: return c[2];   case 3: return c[3];
        case 4: return c[4];   case 5: return c[5];   case 6: return c[6];   case 7: return c[7];
        case 8: return c[8];   case 9: return c[9];   case 10: return c[10]; case 11: return c[11];
        case 12: return c[12]; default: return c[13];
    }
}

template <class T>
__device__ inline T prm(const float* f, const double* d, int i) {
    if constexpr (sizeof(T) == 8) return d[i];
    else return f[i];
}

// floor(a / b) for 0 <= a < 2^20 and 1 <= b < 2^20, from an approximate reciprocal rb of b: the estimate
// is within one of the quotient, and one integer step corrects it
__device__ inline int div_small(int a, int b, float rb) {
    int q = (int)(((float)a + 0.5f) * rb);
    const int r = a - (int)__umul24(q, b);   // q, b < 2^24: the full-rate 24-bit product
    q += (r >= b) - (r < 0);
    return q;
}

// a slot's 4 T values (unit vector and 1/|d|, or a tangent record) in one or two 16-byte LDS reads
template <class T>
__device__ inline void ld4(const T* p, T& a, T& b, T& c, T& d) {
    if constexpr (sizeof(T) == 4) {
        const float4 v = *(const float4*)p;
        a = v.x; b = v.y; c = v.z; d = v.w;
    } else {
        const double2 v0 = *(const double2*)p, v1 = *(const double2*)(p + 2);
        a = v0.x; b = v0.y; c = v1.x; d = v1.y;
    }
}

// Chunk length C (1 .. 64) of an angular block with m rows of at most Dmax pairs: the one that minimises
// rounds (C + w/8), rounds = ceil(m ceil(Dmax / C) / BT) passes of the BT threads over the units and w/8 the
// cost of a chunk's set-up and flush relative to one pair (w from the autotuner, 3 by default; ties: the
// longer chunk).  Lane l tries C = l + 1; every wave finds the same minimum, so C is uniform.
template <int BT>
__device__ inline int pick_chunk(int m, int Dmax, int lane, int w) {
    const int C = min(lane + 1, max(Dmax, 1));
    const int nq = div_small(Dmax + C - 1, C, __builtin_amdgcn_rcpf((float)C));
    const int rounds = ((int)__umul24(m, nq) + BT - 1) / BT;   // m nq < 2^24
    const int key = wave_min_i((int)__umul24(rounds, 8 * C + w) * 64 + (64 - C));
    return 64 - (key & 63);
}

// the 2 NG angular factors of a slot (E then Ed; NG = 1, 2 or 4) in 8- or 16-byte LDS reads
template <class T, int NG>
__device__ inline void ld_e(const T* p, T* e, T* ed) {
    if constexpr (sizeof(T) == 4 && NG == 1) {
        const float2 v = *(const float2*)p;
        e[0] = v.x; ed[0] = v.y;
    } else if constexpr (sizeof(T) == 4 && NG == 2) {
        const float4 v = *(const float4*)p;
        e[0] = v.x; e[1] = v.y; ed[0] = v.z; ed[1] = v.w;
    } else if constexpr (sizeof(T) == 4) {
        const float4 a = *(const float4*)p, b = *(const float4*)(p + 4);
        e[0] = a.x; e[1] = a.y; e[2] = a.z; e[3] = a.w;
        ed[0] = b.x; ed[1] = b.y; ed[2] = b.z; ed[3] = b.w;
    } else {
#pragma unroll
        for (int h = 0; h < NG; ++h) { e[h] = p[h]; ed[h] = p[NG + h]; }
    }
}

// Level bands of the canonical blocks.  SPL = 0: every eta group g uses every level k < NZ.  SPL > 0 (two
// groups): group 0 uses the levels k < SPL and group 1 the levels SPL <= k < NZ -- a model whose small
// zetas come with one eta and large zetas with the other -- so only NZ of the 2 NZ (level, group) products
// are formed per sign.  NA sums per sign; sum c of a sign is (level lev(c), group grp(c)).
template <int NG, int NZ, int SPL>
struct Bands {
    static_assert(SPL == 0 || (NG == 2 && SPL > 0 && SPL < NZ), "a level split takes two eta groups");
    static constexpr int NA = SPL > 0 ? NZ : NZ * NG;
    static constexpr int lo(int g) { return (SPL > 0 && g == 1) ? SPL : 0; }
    static constexpr int hi(int g) { return (SPL > 0 && g == 0) ? SPL : NZ; }   // one past the last level
    static constexpr bool on(int k, int g) { return k >= lo(g) && k < hi(g); }
    static constexpr int lev(int c) { return SPL > 0 ? c : c / NG; }
    static constexpr int grp(int c) { return SPL > 0 ? (c >= SPL ? 1 : 0) : c % NG; }
};

// One angular block whose functions are all canonical: lambda = +-1 and zeta = 2^k, k < NZ <= 7, one
// function per (sign s, level k, eta group g).  Per triplet the windows b_s = |1/2 + s cos/2| are raised
// to 2^k (cp) and 2^k - 1 (cq) by squaring, and the block's sums are the outer products
//   G[s][k][g] += b_s^(2^k) E_j[g] E_k[g]      ((k, g) in the level bands of SPL)
// held in registers with static indices; cidx maps each function of the block to (s*7 + k)*NG + g.
// NZ (a compile-time bound on the levels in use) keeps the loop free of branches; unused levels are
// computed and discarded.  E / Ed of slot q are E[q][g] and E[q][NG + g], zero for g >= ng.
//
// Triplets: every summand is symmetric in its two neighbours, so a block is a set of slot pairs.  A pair is
// (f, v) with the fixed slot f = x of the smaller species side (x < m) and v = (x + d) mod M on the other
// (M its size): a rectangle J x K takes every d < M; the triangle j < k of J x J takes d = 1 .. (n - 1) / 2
// for every x and, for even n, d = n / 2 for x < n / 2 (each unordered pair once).  A row (x, all its d) is
// cut into chunks of C consecutive d, and the workgroup's threads take the units (x, chunk) in turn: f's
// record stays in registers over a chunk, and in the adjoint its dE/dd builds up there and reaches the
// LDS accumulator once per chunk; pick_chunk sets C.
// Each lane's sums (modes 0 and 2) and their wave reduce-scatter are T; the sum over the waves is FP64.
// The adjoint (mode 1) adds each pair's derivative for v into the wave's FP64 accumulators.
template <class T, int NG, int NZ, int MODE, int SPL>
__device__ inline void angular_canon(const BasisLds<T>& L, int bJ, int bK, int nJ, int nK, bool same,
                                     const unsigned char* __restrict__ cidx, const float* wtab, int nA, int goff,
                                     double* gw,
                                     int wv, int lane, int rs_f, bool rs_writer, int chunk_w) {
    using B = Bands<NG, NZ, SPL>;
    constexpr int ADJ = (MODE == 1);
    constexpr int TAN = (MODE == 2);
    constexpr int BT = 64;   // one wave per block
    constexpr int SE = 2 * NG;
    const int tid = lane;
    T acc[2][NZ][NG];
    // adjoint: dE/dG weights w and the weights u = w s 2^(k-1) of d b^(2^k)/d cos = s 2^(k-1) b^(2^k - 1)
    T wgt[2][NZ][NG], wgd[2][NZ][NG];
#pragma unroll
    for (int s = 0; s < 2; ++s)
#pragma unroll
        for (int k = 0; k < NZ; ++k)
#pragma unroll
            for (int g = 0; g < NG; ++g) {
                acc[s][k][g] = T(0);
                wgt[s][k][g] = T(0);
                wgd[s][k][g] = T(0);
            }
    if (ADJ) {
        // the block's dE/dG weights in (s, k, g) order, from the table built before the angular loop
#pragma unroll
        for (int s = 0; s < 2; ++s)
#pragma unroll
            for (int k = 0; k < NZ; ++k)
#pragma unroll
                for (int g = 0; g < NG; ++g) {
                    wgt[s][k][g] = (T)wtab[(s * 7 + k) * NG + g];
                    wgd[s][k][g] = wgt[s][k][g] * (T(1 << k) * (s ? T(-0.5) : T(0.5)));
                }
    }
    // the block's pairs: fixed slots bF + x (x < m), partners bV + y
    const int m = same ? nJ : min(nJ, nK), Mx = same ? nJ : max(nJ, nK);
    const bool sw = !same && nJ > nK;
    const int bF = sw ? bK : bJ, bV = sw ? bJ : bK;
    const int d0 = same ? 1 : 0;
    const int Dmax = same ? nJ / 2 : Mx;                   // longest row
    const int xs = (same && !(nJ & 1)) ? nJ / 2 : m;       // rows x >= xs are one shorter
    const float rm = __builtin_amdgcn_rcpf((float)m);
    const int qB = div_small(BT, m, rm), rB = BT - qB * m; // unit u -> u + BT: chunk + qB, row + rB
    const int C = __builtin_amdgcn_readfirstlane(pick_chunk<BT>(m, Dmax, lane, chunk_w));
    const int nq = div_small(Dmax + C - 1, C, __builtin_amdgcn_rcpf((float)C));
    const int U = m * nq;
    int q = div_small(tid, m, rm), x = tid - (int)__umul24(q, m);   // unit u = q m + x
    // partner records and accumulators from the side's first slot, indexed by y < Mx
    const T* uV = L.u + 4 * bV;
    const T* eV = L.E + SE * bV;
    const T* tV = L.te + 4 * bV;
    double* gV = gw + 3 * bV;
    for (int u = tid; u < U; u += BT) {
        const int j0 = q * C;
        const int cnt = Dmax - (x >= xs ? 1 : 0) - j0;    // pairs of the row from this chunk on
        const int sf = bF + x;
        T fx, fy, fz, fi, Ef[NG], Edf[NG];
        ld4<T>(L.u + 4 * sf, fx, fy, fz, fi);
        ld_e<T, NG>(L.E + SE * sf, Ef, Edf);
        T tfx = T(0), tfy = T(0), tfz = T(0), tfp = T(0);
        if (TAN) ld4<T>(L.te + 4 * sf, tfx, tfy, tfz, tfp);
        T gfx = T(0), gfy = T(0), gfz = T(0);
        unsigned y = x + d0 + j0;
        y = min(y, y - (unsigned)Mx);   // y mod Mx for y < 2 Mx: the difference wraps above y when y < Mx
        for (int j = 0; j < C; ++j) {
            if (j < cnt) {
                T vx, vy, vz, vi, Ev[NG], Edv[NG];
                ld4<T>(uV + 4 * y, vx, vy, vz, vi);
                ld_e<T, NG>(eV + SE * y, Ev, Edv);
                const T cs = fx * vx + fy * vy + fz * vz;
                T D[NG];
#pragma unroll
                for (int g = 0; g < NG; ++g) D[g] = Ef[g] * Ev[g];
                // cp[s][k] = b_s^(2^k), cq[s][k] = b_s^(2^k - 1) (tangent); the levels no band reads are
                // dropped by the compiler
                T cp[2][NZ], cq[2][NZ];
                cp[0][0] = fabs(T(0.5) + T(0.5) * cs);
                cp[1][0] = fabs(T(0.5) - T(0.5) * cs);
                cq[0][0] = T(1);
                cq[1][0] = T(1);
#pragma unroll
                for (int k = 1; k < NZ; ++k)
#pragma unroll
                    for (int s = 0; s < 2; ++s) {
                        cq[s][k] = TAN ? ((k == 1) ? cp[s][0] : cq[s][k - 1] * cp[s][k - 1]) : T(0);
                        cp[s][k] = cp[s][k - 1] * cp[s][k - 1];
                    }
                if (MODE == 0) {
#pragma unroll
                    for (int k = 0; k < NZ; ++k)
#pragma unroll
                        for (int s = 0; s < 2; ++s)
#pragma unroll
                            for (int g = 0; g < NG; ++g)
                                if (B::on(k, g)) acc[s][k][g] += cp[s][k] * D[g];
                } else if (TAN) {
                    T tvx, tvy, tvz, tvp;
                    ld4<T>(tV + 4 * y, tvx, tvy, tvz, tvp);
                    const T dcs = (vx * tfx + vy * tfy + vz * tfz - cs * tfp) * fi +
                                  (fx * tvx + fy * tvy + fz * tvz - cs * tvp) * vi;
                    T X[NG], Y[NG];
#pragma unroll
                    for (int g = 0; g < NG; ++g) {
                        X[g] = D[g] * dcs;
                        Y[g] = Edf[g] * Ev[g] * tfp + Ef[g] * Edv[g] * tvp;
                    }
#pragma unroll
                    for (int k = 0; k < NZ; ++k)
#pragma unroll
                        for (int s = 0; s < 2; ++s) {
                            const T dp = T(1 << k) * (s ? T(-0.5) : T(0.5)) * cq[s][k];   // d b^(2^k) / d cos
#pragma unroll
                            for (int g = 0; g < NG; ++g)
                                if (B::on(k, g)) acc[s][k][g] += dp * X[g] + cp[s][k] * Y[g];
                        }
                } else {
                    // Over the levels lo <= k < hi of group g's band:
                    // S_g = sum_s b^(2^lo) (w_lo + b^(2^lo) (w_lo+1 + b^(2^(lo+1)) (...))) = sum_s,k w_k b^(2^k) and
                    // Sp_g = sum_s b^(2^lo - 1) (u_lo + b^(2^lo) (u_lo+1 + ...)) = sum_s,k u_k b^(2^k - 1),
                    // u_k = w_k d b^(2^k)/d cos / b^(2^k - 1): Horner over the squares b^(2^k)
                    T S[NG], Sp[NG];
#pragma unroll
                    for (int s = 0; s < 2; ++s)
#pragma unroll
                        for (int g = 0; g < NG; ++g) {
                            T h = T(0), hd = T(0), q = T(1);
#pragma unroll
                            for (int k = NZ - 1; k >= 0; --k) {
                                if (k == B::hi(g) - 1) {
                                    h = wgt[s][k][g];
                                    hd = wgd[s][k][g];
                                } else if (B::on(k, g)) {
                                    h = h * cp[s][k] + wgt[s][k][g];
                                    hd = hd * cp[s][k] + wgd[s][k][g];
                                }
                                if (k < B::lo(g)) q = (k == B::lo(g) - 1) ? cp[s][k] : q * cp[s][k];   // b^(2^lo - 1)
                            }
                            const T hs = h * cp[s][B::lo(g)], hds = (B::lo(g) > 0) ? hd * q : hd;
                            S[g] = s ? S[g] + hs : hs;
                            Sp[g] = s ? Sp[g] + hds : hds;
                        }
                    T dEdc = Sp[0] * D[0], dEdf = S[0] * (Edf[0] * Ev[0]), dEdv = S[0] * (Ef[0] * Edv[0]);
#pragma unroll
                    for (int g = 1; g < NG; ++g) {
                        dEdc += Sp[g] * D[g];
                        dEdf += S[g] * (Edf[g] * Ev[g]);
                        dEdv += S[g] * (Ef[g] * Edv[g]);
                    }
                    const T cc = dEdc * cs;
                    const T kf = dEdf - cc * fi, kv = dEdv - cc * vi;
                    const T xf = dEdc * fi, xv = dEdc * vi;
                    gfx = fma(xf, vx, fma(kf, fx, gfx));
                    gfy = fma(xf, vy, fma(kf, fy, gfy));
                    gfz = fma(xf, vz, fma(kf, fz, gfz));
                    double* gv = (double*)((char*)gV + __umul24(y, 24u));   // 24-bit product: one full-rate op
                    atomicAdd(gv, (double)(kv * vx + xv * fx));
                    atomicAdd(gv + 1, (double)(kv * vy + xv * fy));
                    atomicAdd(gv + 2, (double)(kv * vz + xv * fz));
                }
            }
            ++y;
            y = min(y, y - (unsigned)Mx);
        }
        if (ADJ && cnt > 0) {
            double* gf = gw + 3 * sf;
            atomicAdd(gf, (double)gfx);
            atomicAdd(gf + 1, (double)gfy);
            atomicAdd(gf + 2, (double)gfz);
        }
        x += rB;
        q += qB;
        if (x >= m) {
            x -= m;
            ++q;
        }
    }
    if (!ADJ) {
        // wave reduce-scatter of the 2 NA sums; the lane that writes sum f puts it in its (sign, level, group)
        // slot
        constexpr int M0 = 2 * B::NA;
        static_assert(M0 <= 64, "one sum per lane at most");
        T v[M0];
#pragma unroll
        for (int s = 0; s < 2; ++s)
#pragma unroll
            for (int c = 0; c < B::NA; ++c) v[s * B::NA + c] = acc[s][B::lev(c)][B::grp(c)];
        rs_any<5, M0, T>(v, lane);
        const double tot = (double)v[0];
        if (rs_writer) {
            const int f = rs_f;
            const int s = f / B::NA, c = f - s * B::NA;
            L.red[wv * 64 + (s * 7 + B::lev(c)) * NG + B::grp(c)] = tot;
        }
        wave_sync();
        for (int m = lane; m < nA; m += 64) L.row[goff + m] = L.red[wv * 64 + cidx[m]];
        wave_sync();   // the wave's next block reuses its red slots
    }
}

// The wave that takes the next block of a phase: the least loaded one (the first on ties), charged with the
// block's cost.  Every wave runs the same deal over the same blocks, so the owners agree without a barrier.
template <int BW>
__device__ inline int deal(int* load, int cost) {
    int w = 0;
#pragma unroll
    for (int k = 1; k < BW; ++k) w = (load[k] < load[w]) ? k : w;
#pragma unroll
    for (int k = 0; k < BW; ++k) load[k] += (k == w) ? cost : 0;
    return w;
}

// Radial and angular sums of one atom (slots, tables and row already in LDS).  PM 0: G into row;
// PM 1: per-slot dE/dd into the per-wave accumulators g from the dE/dG weights in row; PM 2: tangent
// rows into row.  Ends with a workgroup barrier.
template <class T, int NG, int NZ, int PM, int GEN, int BW, int SPL, bool LEAN = false>
__device__ inline void basis_sums(const BasisLds<T>& L, const DevSpecies& sp, const DevBasisParams& P,
                                  const int* s_cnt, int nt, int nbm, int tid, int wv, int lane) {
    constexpr int MODE = PM;
    constexpr int BT = BW * 64;
    constexpr int ADJ = (PM == 1);
    constexpr int TAN = (PM == 2);
    constexpr int MAXA = 16;   // functions per register chunk of the generic path
    static_assert(!LEAN || (ADJ && !GEN), "the lean layout is the split adjoint's without the generic path");
    constexpr bool ROWF = LEAN && sizeof(T) == 4;   // basis_rowf: FP32 dE/dG row
    const int ng = sp.ng;
    if (LEAN && wv > 0) {
        // waves 1.. clear their own accumulators (wave 1's held |d| until the per-slot pass); the barriers
        // below order this before any wave adds into them
        double* gw0 = L.g + wv * 3 * nbm;
        for (int k = lane; k < 3 * s_cnt[nt]; k += 64) gw0[k] = 0.0;
    }
    // --- radial blocks (the lean adjoint added them in the per-slot pass): one wave per neighbour species J.  Sums over neighbours (G, tangent rows): lane =
    // function, lane group = neighbour subset, the lanes of a function sum their groups in FP64.  Sums over
    // functions (the adjoint's dE/dr per neighbour): lane = (neighbour, function group), the neighbour's
    // |d| and cutoff values in registers, the groups summed across lanes once per neighbour ------------
    if (!LEAN) {
        int load[BW];
#pragma unroll
        for (int w = 0; w < BW; ++w) load[w] = 0;
        for (int J = 0; J < nt; ++J) {
            const int bJ = s_cnt[J], nJ = s_cnt[J + 1] - s_cnt[J];
            const int nR = sp.r_n[J];
            if (nJ == 0 || nR == 0) continue;
            if (deal<BW>(load, nJ * nR) != wv) continue;
            const int form = sp.r_form[J];
            const int goff = sp.r_goff[J], poff = sp.r_poff[J];
            if (ADJ) {
                int ns = 1;   // lanes per function group: neighbours side by side, at most 64
                while (ns < nJ && ns < 64) ns <<= 1;
                const int fg = 64 / ns, h = lane / ns;
                const int rq0 = poff - sp.r_poff[0];
                for (int q0 = 0; q0 < nJ; q0 += ns) {
                    const int q = q0 + lane % ns;
                    const bool on = q < nJ;
                    const int slot = bJ + (on ? q : 0);
                    const double r = slot_r64(L, slot);
                    const T fc = L.fR[slot], gc = L.gR[slot];
                    T acc = T(0);
                    // the form is uniform: the Gaussian (every benchmark model) without the dispatch per function
                    if (on && form == ACE_RAD_GAUSSIAN)
                        for (int n = h; n < nR; n += fg) {
                            T v, dv;
                            t_radial_r64<T>(ACE_RAD_GAUSSIAN, L.rp[rq0 + n], L.rp[sp.nR + rq0 + n], r, fc, gc, v, dv);
                            acc += (T)L.row[goff + n] * dv;
                        }
                    else if (on)
                        for (int n = h; n < nR; n += fg) {
                            T v, dv;
                            t_radial_r64<T>(form, L.rp[rq0 + n], L.rp[sp.nR + rq0 + n], r, fc, gc, v, dv);
                            acc += (T)L.row[goff + n] * dv;
                        }
                    acc = t_cross_sum<T>(acc, ns);
                    if (h == 0 && on) L.fr[slot] += acc;
                }
                continue;
            }
            for (int n0 = 0; n0 < nR; n0 += 64) {
                const int nc = min(64, nR - n0);
                int nrp = 1;
                while (nrp < nc) nrp <<= 1;
                const int gpw = 64 / nrp;
                const int n = lane % nrp, grp = lane / nrp;
                const bool on = n < nc;
                const int rq = poff - sp.r_poff[0] + n0 + n;
                const double eta = on ? L.rp[rq] : 0.0;
                const double rs = on ? L.rp[sp.nR + rq] : 0.0;
                T acc = T(0);
                const int rounds = (nJ + gpw - 1) / gpw;
                if (form == ACE_RAD_GAUSSIAN)
                    for (int t = 0; t < rounds; ++t) {
                        const int q = grp + t * gpw;
                        if (on && q < nJ) {
                            const int slot = bJ + q;
                            T v, dv;
                            t_radial_r64<T>(ACE_RAD_GAUSSIAN, eta, rs, slot_r64(L, slot), L.fR[slot], L.gR[slot], v, dv);
                            acc += TAN ? dv * L.te[4 * slot + 3] : v;
                        }
                    }
                else
                    for (int t = 0; t < rounds; ++t) {
                        const int q = grp + t * gpw;
                        if (on && q < nJ) {
                            const int slot = bJ + q;
                            T v, dv;
                            t_radial_r64<T>(form, eta, rs, slot_r64(L, slot), L.fR[slot], L.gR[slot], v, dv);
                            acc += TAN ? dv * L.te[4 * slot + 3] : v;
                        }
                    }
                const double tot = t_cross_sum<double>((double)acc, nrp);
                if (lane < nc) L.row[goff + n0 + lane] = tot;
            }
        }
    }
    if (ADJ) {
        // dE/dG of every canonical block scattered to its (sign, level, group) table in red: two barriers
        // for all blocks
        const int nab = nt * (nt + 1) / 2;
        float* wt = (float*)L.red;
        for (int q = tid; q < nab * 14 * NG; q += BT) wt[q] = 0.f;
        __syncthreads();
        for (int b = 0; b < nab; ++b)
            if (!GEN || sp.a_canon[b]) {
                const int o = sp.a_poff[b] - sp.a_poff[0];
                for (int m = tid; m < sp.a_n[b]; m += BT)
                    wt[14 * NG * b + L.ci[o + m]] = ROWF ? L.rowf[sp.a_goff[b] + m] : (float)L.row[sp.a_goff[b] + m];
            }
        __syncthreads();
    }

    // --- angular blocks (J <= K): canonical ones one wave each, then the others with all waves ----------
    double* gw = L.g + wv * 3 * nbm;
    int rs_f = 0;
    bool rs_writer = false;
    if (!ADJ) rs_any_owner<2 * Bands<NG, NZ, SPL>::NA>(lane, rs_f, rs_writer);
    {
        int load[BW];
#pragma unroll
        for (int w = 0; w < BW; ++w) load[w] = 0;
        for (int K = 0; K < nt; ++K) {
            for (int J = 0; J <= K; ++J) {
                const int b = K * (K + 1) / 2 + J;
                const int nA = sp.a_n[b];
                const bool same = (J == K);
                const int bJ = s_cnt[J], bK = s_cnt[K];
                const int nJ = s_cnt[J + 1] - bJ, nK = s_cnt[K + 1] - bK;
                const int T3 = same ? nJ * (nJ - 1) / 2 : nJ * nK;
                if (nA == 0 || T3 == 0 || (GEN && !sp.a_canon[b])) continue;
                if (deal<BW>(load, T3) != wv) continue;
                angular_canon<T, NG, NZ, MODE, SPL>(L, bJ, bK, nJ, nK, same, L.ci + (sp.a_poff[b] - sp.a_poff[0]),
                                               (const float*)L.red + 14 * NG * b, nA, sp.a_goff[b], gw, wv, lane,
                                               rs_f, rs_writer, P.chunk_w);
            }
        }
    }
    for (int K = 0; K < (GEN ? nt : 0); ++K) {
        for (int J = 0; J <= K; ++J) {
            const int b = K * (K + 1) / 2 + J;
            const int nA = sp.a_n[b];
            const bool same = (J == K);
            const int bJ = s_cnt[J], bK = s_cnt[K];
            const int nJ = s_cnt[J + 1] - bJ, nK = s_cnt[K + 1] - bK;
            const int T3 = same ? nJ * (nJ - 1) / 2 : nJ * nK;
            if (nA == 0 || T3 == 0 || sp.a_canon[b]) continue;
            const int goff = sp.a_goff[b], poff = sp.a_poff[b], aform = sp.a_form[b];
            const float rnK = 1.f / (float)(same ? nJ : min(nJ, nK));
            for (int m0 = 0; m0 < nA; m0 += MAXA) {
                T lamf[MAXA], zeta[MAXA], acc[MAXA], w[MAXA];
                int zint[MAXA], grp[MAXA], sel[MAXA];
                int zmax = 1;
#pragma unroll
                for (int m = 0; m < MAXA; ++m) {
                    const bool on = m0 + m < nA;
                    const int fi = poff + m0 + m;
                    lamf[m] = on ? prm<T>(P.a_lamf, P.a_lamf64, fi) : T(0.5);
                    zeta[m] = on ? prm<T>(P.a_zeta, P.a_zeta64, fi) : T(1);
                    zint[m] = on ? P.a_zint[fi] : 1;
                    grp[m] = on ? P.a_grp[fi] : 0;
                    sel[m] = on ? P.a_sel[fi] : 0;
                    acc[m] = T(0);
                    w[m] = (ADJ && on) ? (T)L.row[goff + m0 + m] : T(0);
                    zmax = max(zmax, zint[m]);
                }
                int nk = 1;
                while (nk < 7 && (1 << nk) <= zmax) ++nk;
                for (int t = tid; t < T3; t += BT) {
                    int a, c;
                    tri_decode(t, same, nJ, nK, rnK, a, c);
                    const int sa = bJ + a, sc = bK + c;
                    const T uax = L.u[4 * sa], uay = L.u[4 * sa + 1], uaz = L.u[4 * sa + 2];
                    const T ucx = L.u[4 * sc], ucy = L.u[4 * sc + 1], ucz = L.u[4 * sc + 2];
                    const T cs = uax * ucx + uay * ucy + uaz * ucz;
                    // distance factors: X = Da Dc, dX/d|d_a| = Dda Dc, dX/d|d_c| = Da Ddc
                    T Da[ACE_MAXG], Dc[ACE_MAXG], Dda[ACE_MAXG], Ddc[ACE_MAXG];
                    if (aform == ACE_ANG_GAUSS) {
#pragma unroll
                        for (int g = 0; g < ACE_MAXG; ++g) {
                            const bool on = g < ng;
                            Da[g] = on ? L.E[sa * 2 * NG + g] : T(0);
                            Dc[g] = on ? L.E[sc * 2 * NG + g] : T(0);
                            Dda[g] = (on && (ADJ || TAN)) ? L.E[sa * 2 * NG + NG + g] : T(0);
                            Ddc[g] = (on && (ADJ || TAN)) ? L.E[sc * 2 * NG + NG + g] : T(0);
                        }
                    } else {
                        // not separable in the two distances: F(r_a^2 + r_c^2) per triplet, in FP64
                        const double ra = slot_r64(L, sa), rc = slot_r64(L, sc), r2s = ra * ra + rc * rc;
                        const T fa = L.fA[sa], fc = L.fA[sc], ga = L.gA[sa], gc = L.gA[sc];
#pragma unroll
                        for (int g = 0; g < ACE_MAXG; ++g) {
                            double F = 0.0, q = 0.0;
                            if (g < ng) angular_factor(aform, sp.ietar2_64[g], r2s, F, q);
                            Da[g] = (g < ng) ? fa : T(0);
                            Dc[g] = fc * (T)F;
                            Dda[g] = ga + (T)(2.0 * ra * q) * fa;
                            Ddc[g] = (T)F * (gc + (T)(2.0 * rc * q) * fc);
                        }
                    }
                    // cp[k + 7s] = b_s^(2^k), cq[k + 7s] = b_s^(2^k - 1), b_0 = (1+cos)/2, b_1 = (1-cos)/2
                    T cp[14], cq[14];
                    cp[0] = fabs(T(0.5) + T(0.5) * cs);
                    cp[7] = fabs(T(0.5) - T(0.5) * cs);
                    cq[0] = T(1);
                    cq[7] = T(1);
#pragma unroll
                    for (int k = 1; k < 7; ++k) {
                        const bool on = k < nk;
                        cq[k] = on ? cq[k - 1] * cp[k - 1] : T(0);
                        cp[k] = on ? cp[k - 1] * cp[k - 1] : T(0);
                        cq[k + 7] = on ? cq[k + 6] * cp[k + 6] : T(0);
                        cp[k + 7] = on ? cp[k + 6] * cp[k + 6] : T(0);
                    }
                    T X[ACE_MAXG], Y[ACE_MAXG], S[ACE_MAXG], Sp[ACE_MAXG];
#pragma unroll
                    for (int g = 0; g < ACE_MAXG; ++g) {
                        X[g] = Da[g] * Dc[g];
                        Y[g] = T(0);
                        S[g] = T(0);
                        Sp[g] = T(0);
                    }
                    if (TAN) {
                        const T tax = L.te[4 * sa], tay = L.te[4 * sa + 1], taz = L.te[4 * sa + 2], tap = L.te[4 * sa + 3];
                        const T tcx = L.te[4 * sc], tcy = L.te[4 * sc + 1], tcz = L.te[4 * sc + 2], tcp = L.te[4 * sc + 3];
                        const T dcs = (ucx * tax + ucy * tay + ucz * taz - cs * tap) * L.u[4 * sa + 3] +
                                      (uax * tcx + uay * tcy + uaz * tcz - cs * tcp) * L.u[4 * sc + 3];
#pragma unroll
                        for (int g = 0; g < ACE_MAXG; ++g) {
                            Y[g] = Dda[g] * Dc[g] * tap + Da[g] * Ddc[g] * tcp;
                            X[g] *= dcs;
                        }
                    }
#pragma unroll
                    for (int m = 0; m < MAXA; ++m) {
                        if (m0 + m < nA) {
                            T base, p1;          // window and window^(zeta-1)
                            if (sel[m] >= 0) {
                                base = (sel[m] < 7) ? cp[0] : cp[7];
                                p1 = sel14<T>(cq, sel[m]);
                            } else {
                                base = fabs(T(0.5) + lamf[m] * cs);
                                p1 = (zint[m] >= 1) ? t_pow_window<T>(base, T(0), zint[m] - 1)
                                                    : t_pow_window<T>(base, zeta[m] - T(1), -1);
                            }
                            if (MODE == 0) {
                                acc[m] += p1 * base * pick4<T>(X, grp[m]);
                            } else if (TAN) {
                                acc[m] += zeta[m] * lamf[m] * p1 * pick4<T>(X, grp[m]) + p1 * base * pick4<T>(Y, grp[m]);
                            } else {
                                const T s = w[m] * p1 * base, s1 = w[m] * zeta[m] * lamf[m] * p1;
#pragma unroll
                                for (int gg = 0; gg < ACE_MAXG; ++gg) {
                                    S[gg] += (grp[m] == gg) ? s : T(0);
                                    Sp[gg] += (grp[m] == gg) ? s1 : T(0);
                                }
                            }
                        }
                    }
                    if (ADJ) {
                        T dEdc = T(0), dEda = T(0), dEdcc = T(0);
#pragma unroll
                        for (int g = 0; g < ACE_MAXG; ++g) {
                            dEdc += Sp[g] * X[g];
                            dEda += S[g] * Dda[g] * Dc[g];
                            dEdcc += S[g] * Da[g] * Ddc[g];
                        }
                        const T ia = L.u[4 * sa + 3], ic = L.u[4 * sc + 3];
                        const T ka = dEda - dEdc * cs * ia, kc = dEdcc - dEdc * cs * ic;
                        const T xa = dEdc * ia, xc = dEdc * ic;
                        atomicAdd(&gw[3 * sa], (double)(ka * uax + xa * ucx));
                        atomicAdd(&gw[3 * sa + 1], (double)(ka * uay + xa * ucy));
                        atomicAdd(&gw[3 * sa + 2], (double)(ka * uaz + xa * ucz));
                        atomicAdd(&gw[3 * sc], (double)(kc * ucx + xc * uax));
                        atomicAdd(&gw[3 * sc + 1], (double)(kc * ucy + xc * uay));
                        atomicAdd(&gw[3 * sc + 2], (double)(kc * ucz + xc * uaz));
                    }
                }
                if (!ADJ) {
#pragma unroll
                    for (int m = 0; m < MAXA; ++m) {
                        if (m0 + m < nA) {
                            const double t = (double)t_wave_sum<T>(acc[m]);
                            if (lane == 0) L.red[wv * 64 + m] = t;
                        }
                    }
                    __syncthreads();
                    if (tid < MAXA && m0 + tid < nA) {
                        double s = 0.0;
                        for (int ww = 0; ww < BW; ++ww) s += L.red[ww * 64 + tid];
                        L.row[goff + m0 + tid] = s;
                    }
                    __syncthreads();
                }
            }
        }
    }
    __syncthreads();
}

// The atom's network in wave 0 of the workgroup (energy and dE/dG, FP32; ANN::fpbp and DODZ row 0,
// ml/nn.cpp:1402-1420, 1836-1862).  On entry row holds G; on exit e_out holds the atomic energy (with E0)
// and row holds dE/dG (the input scaling inpw folded in).  The species' parameters are in L.nw in the
// matrix-core layout (W_l as [k][o], row stride mlp_ld(w_{l+1}), biases at bl_off); after them: per layer
// the activations and derivatives.  A chunk of a layer's outputs is spread over the 64 lanes, G lanes per
// output (G a power of two); a lane's strided partial sum runs in eight accumulators that are combined, and
// then summed over the group by DPP moves, in a fixed order.  One wave needs no workgroup barrier between
// layers (LDS accesses of a wave complete in order; wave_sync keeps the compiler from reordering them);
// one barrier at the end publishes dE/dG to the other waves.
__device__ inline int net_group(int nout) {
    int g = 1;
    while (g < 64 && 2 * g * nout <= 64) g <<= 1;
    return g;
}

// sum_{k = r, r + G, ... < n} w[k * ldw] x[k] in eight independent accumulators (16 LDS loads in flight per
// round trip), combined in a fixed order
template <class X>
__device__ inline float net_dot(const float* w, int ldw, const X* x, int r, int G, int n) {
    float p[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) p[u] = 0.f;
    int k = r;
    for (; k + 7 * G < n; k += 8 * G) {
        float wv[8], xv[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            wv[u] = w[(k + u * G) * ldw];
            xv[u] = (float)x[k + u * G];
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) p[u] += wv[u] * xv[u];
    }
    for (int u = 0; k < n; k += G, ++u) p[u & 7] += w[k * ldw] * (float)x[k];
    return ((p[0] + p[1]) + (p[2] + p[3])) + ((p[4] + p[5]) + (p[6] + p[7]));
}

template <class T>
__device__ inline void block_network(const BasisLds<T>& L, const DevSpecies& sp, int tid, double* e_out) {
    if (tid < 64) {
        const int lane = tid;
        const int K = sp.K, NL = sp.L;
        const float* W = L.nw;
        float* act = L.nw + sp.wp_n;
        for (int k = lane; k < K; k += 64) L.row[k] = L.sc[k] * (L.row[k] + L.sc[K + k]);   // inpw (G + inpb), FP64
        wave_sync();
        // forward
        int ain = -1, aoff = 0;
        for (int l = 0; l < NL; ++l) {
            const int ni = sp.width[l], no = sp.width[l + 1], ldw = mlp_ld(no), wp = (no + 3) & ~3;
            const float* Wl = W + sp.wl_off[l];
            const float* bl = W + sp.bl_off[l];
            float* A = act + aoff;
            float* D = A + wp;
            const int fa = (l == NL - 1) ? ACE_ACT_LINEAR : sp.act;
            for (int o0 = 0; o0 < no; o0 += 64) {
                const int nout = min(64, no - o0), G = net_group(nout);
                const int o = o0 + lane / G, r = lane % G;
                const bool on = o < o0 + nout;
                const int oc = on ? o : o0;
                float part = (ain < 0) ? net_dot(Wl + oc, ldw, L.row, r, G, ni) : net_dot(Wl + oc, ldw, act + ain, r, G, ni);
                part = on ? part : 0.f;
                part = t_group_sum<float>(part, G);
                if (r == 0 && on) {
                    float a, d;
                    act_fd(fa, sp.c, part + bl[o], a, d);
                    A[o] = a;
                    D[o] = d;
                }
            }
            wave_sync();
            ain = aoff;
            aoff += 2 * wp;
        }
        // energy (the last layer has width 1) and the output delta
        float* A = act + ain;
        const int wpL = (sp.width[NL] + 3) & ~3;
        if (lane == 0) {
            *e_out = (double)(sp.outb + sp.outw * A[0]) + sp.E0;
            A[0] = sp.outw * A[wpL];   // delta = outw * f'(z), f linear
        }
        wave_sync();
        // backward: delta_{l-1}[k] = (sum_o W_l[k][o] delta_l[o]) f'(z_{l-1})[k]; layer 0 gives dE/dG
        int dl = ain;   // offset of delta_l (overwrites the activations of layer l)
        for (int l = NL - 1; l >= 0; --l) {
            const int ni = sp.width[l], no = sp.width[l + 1], ldw = mlp_ld(no);
            const float* Wl = W + sp.wl_off[l];
            const int wpi = (ni + 3) & ~3;
            const int dm = (l > 0) ? dl - 2 * wpi : 0;   // offset of layer l-1's activations
            for (int k0 = 0; k0 < ni; k0 += 64) {
                const int nout = min(64, ni - k0), G = net_group(nout);
                const int k = k0 + lane / G, r = lane % G;
                const bool on = k < k0 + nout;
                const int kc = on ? k : k0;
                float part = net_dot(Wl + kc * ldw, 1, act + dl, r, G, no);
                part = on ? part : 0.f;
                part = t_group_sum<float>(part, G);
                if (r == 0 && on) {
                    if (l > 0) act[dm + k] = part * act[dm + wpi + k];
                    else L.row[k] = (double)part * L.sc[k];
                }
            }
            wave_sync();
            dl = dm;
        }
    }
    __syncthreads();
}

// Fixed-point accumulation of forces, energies and virials: 64-bit integer atomics in units of 2^-34
// (resolution 5.8e-11, range 5.4e8) are exact and order independent, so the sums do not depend on the
// order in which workgroups add them.
#define FIX_SCALE 17179869184.0
__device__ inline unsigned long long fix_of(double v) { return (unsigned long long)__double2ll_rn(v * FIX_SCALE); }

// Kernel inputs and outputs.  Neighbours of atom i: counts[i] keys in row keys[i * cap ...], sorted in
// the reference order (key = j_local << 30 | image offset by 512 in three 10-bit fields).  The gradient
// modes add -dE/dr into the fixed-point accumulators F_fix (per atom; or, gathered, leave each pair's
// gradient in gin for k_force_gather), E_fix and W_fix (per structure) and zero the entries of the other
// buffer that this workgroup owns, which the next call uses.
struct BasisIO {
    const int* type;
    const double* pos;
    const int* struc_of;
    const int* atom_off;
    const StrucInfo* sinfo;
    const long long* keys;
    const int* counts;
    int cap;
    int nrp;            // most radial functions of a species (LDS table size)
    const int* row_of;
    int ldg;
    float* G;
    const float* dEdG;
    int* err;
    int lds_bytes;      // the launch's dynamic LDS: a layout that would not fit is refused
    const double* eps;
    const float* wpad;
    double* e_atom;
    unsigned long long *F_fix, *E_fix, *W_fix;
    unsigned long long *F_nxt, *E_nxt, *W_nxt;
    // gradient launches: gin == nullptr adds each pair's g to F_j and -g to F_i with fixed-point atomics;
    // else g of pair (i -> j, n) goes to the entry of its reverse (j -> i, -n) in j's row,
    // gin[3 (j cap + p) ..] (zero for a pair past the cutoff), -sum g to fown[3 i ..], and k_force_gather
    // sums each row
    double* gin;
    double* fown;
    // reverse entries (j cap + p) of the key-row entries (i, e): written by the symmetry-function pass, read
    // by the gradient pass (nullptr: not written / searched in the gradient pass)
    int* rev;
};


// Entry of the reverse pair (j -> i, -n) of key `key` (pair i -> j, n, j = j_local + a0) in j's sorted key
// row: binary search; -1 if absent (the list is symmetric, so only a refused build gets there).
__device__ inline long long reverse_entry(const long long* __restrict__ keys, int cap, const int* __restrict__ counts,
                                         long long key, int il, int a0) {
    const int j = (int)(key >> 30) + a0;
    const int c = (int)(key & 0x3fffffff);
    const long long target = ((long long)il << 30) | ((long long)(1024 - (c >> 20)) << 20) |
                             ((1024 - ((c >> 10) & 1023)) << 10) | (1024 - (c & 1023));
    const int n = min(counts[j], cap);
    const long long* rj = keys + (long long)j * cap;
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (rj[mid] < target) lo = mid + 1;
        else hi = mid;
    }
    return (lo < n && rj[lo] == target) ? (long long)j * cap + lo : -1;
}

// GEN = 0 compiles the kernel for models whose angular blocks are all canonical (no generic path).
// NG: eta groups per species (1, 2 or 4, at least the model's largest count); NZ: canonical power levels
// (3, 4, 5 or 7, at least the model's largest log2(zeta) + 1); SPL: the level split of two eta groups
// (0: none; Bands).
// MODE 0: network input rows; 1: gradients from dE/dG rows; 2: tangent rows; 3: fused energy and forces.
template <class T, int NG, int NZ, int MODE, int GEN, int BW = 4, int SPL = 0>
// The FP32 symmetry-function kernels of up to ten (eta group, level) pairs without the generic path ask
// the compiler for 8 waves per SIMD: their SGPRs (106, 7 waves) then fit with a few spilled to VGPR lanes
// and no VGPR spills (the larger instantiations would spill VGPRs); 4.94 -> 4.46 ms on the 568k-atom batch.
// The FP32 adjoint and tangent kernels take the same hint up to two eta groups and three levels, where they
// spill no VGPR (training step, 64 structures of 20,245 atoms: 0.612 -> 0.591 ms).
__global__ __launch_bounds__(BW * 64) __attribute__((amdgpu_waves_per_eu(
    (sizeof(T) == 4 && GEN == 0 && ((MODE == 0 && NG * NZ <= 10) || ((MODE == 1 || MODE == 2) && NG <= 2 && NZ <= 3))) ? 8 : 1))) void k_basis(int natoms, const DevModel* __restrict__ M, DevBasisParams P, BasisIO io,
                                              int nbm, int nt, int kmax) {
    constexpr int ADJ = (MODE == 1);
    constexpr int TAN = (MODE == 2);
    constexpr int FUSED = (MODE == 3);        // symmetry functions, network and adjoint in one pass
    constexpr int GACC = (ADJ || FUSED);      // per-slot gradient accumulators
    constexpr bool LEAN = (MODE == 1 && !GEN);   // basis_lean: no per-slot radial arrays
    constexpr bool ROWF = LEAN && sizeof(T) == 4;  // basis_rowf: FP32 dE/dG row
    constexpr bool F64 = sizeof(T) == 8;
    constexpr int BT = BW * 64;
    // the waves that stage the radial and the angular tables (wave 1 copies the species record)
    constexpr int W_RAD = BW > 2 ? 2 : 1, W_ANG = BW > 3 ? 3 : BW - 1;
    static_assert(BW >= 2 && BW <= 4, "two to four waves per atom");
    extern __shared__ double4 basis_smem[];
    __shared__ int s_cnt[ACE_MAXT + 1];
    __shared__ DevSpecies s_sp;   // the fused kernel's copy of the species record (LDS does not limit it)
    const int tid = threadIdx.x, wv = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63;
    const int i = blockIdx.x;
    if (i >= natoms) return;
    const int st = io.struc_of[i];
    if (GACC) {
        // the other buffer's entries owned by this atom, for the next call
        if (tid < 3) io.F_nxt[3 * i + tid] = 0ull;
        if (i == io.atom_off[st] && tid < 10) {
            if (tid == 0) io.E_nxt[st] = 0ull;
            else io.W_nxt[9 * st + tid - 1] = 0ull;
        }
    }
    const int I = io.type[i];
    const DevSpecies& gsp = M->sp[I];
    const BasisLds<T> L = basis_lds<T, BW>(basis_smem, nbm, NG, kmax, io.nrp, MODE, nt, GEN != 0);
    if ((long long)((char*)L.nw - (char*)basis_smem) > io.lds_bytes) {
        if (tid == 0) atomicAdd(io.err, 1);
        return;
    }
    const int nn = io.counts[i];
    const long long* krow = io.keys + (long long)i * io.cap;
    const float rcut = gsp.rc_max;
    const int row = io.row_of[i];

    // --- waves 1.. stage the species tables in LDS while wave 0 buckets the neighbours; the split kernels
    // read the species record from memory with uniform (scalar) loads, which saves its LDS for more atoms
    // in flight, the fused one (a round of workgroups or a few, latency-bound) from an LDS copy ----------
    if (FUSED && wv == 1) {
        constexpr int n16 = (int)(sizeof(DevSpecies) / 16);
        static_assert(sizeof(DevSpecies) % 16 == 0, "DevSpecies is copied in 16-byte words");
        static_assert(n16 >= 64 && n16 <= 128, "DevSpecies copy covers 64 to 128 words");
        const int4* src = (const int4*)&gsp;
        int4* dst = (int4*)&s_sp;
        const int4 v0 = src[lane];
        const int4 v1 = (64 + lane < n16) ? src[64 + lane] : v0;
        dst[lane] = v0;
        if (64 + lane < n16) dst[64 + lane] = v1;
    }
    if (wv == W_RAD) {
        const int nR = gsp.nR, r0 = gsp.r_poff[0];
        for (int q = lane; q < nR; q += 64) {
            L.rp[q] = P.r_eta64[r0 + q];
            L.rp[nR + q] = P.r_rs64[r0 + q];
        }
        if (ADJ)
            for (int k = lane; k < gsp.K; k += 64) {
                const float v = io.dEdG[(long long)row * io.ldg + k];
                if (ROWF) L.rowf[k] = v;
                else L.row[k] = (double)v;
            }
    }
    // the symmetry-function pass of a gathered energy/force pass finds, for each entry of the row, the entry
    // of its reverse pair in the neighbour's row (where the gradient pass writes); wave 1 searches while
    // wave 0 buckets
    if (MODE == 0 && wv == 1 && io.rev && nn <= io.cap) {
        const int a0s = io.sinfo[st].atom_off;
        for (int e = lane; e < nn; e += 64)
            io.rev[(long long)i * io.cap + e] = (int)reverse_entry(io.keys, io.cap, io.counts, krow[e], i - a0s, a0s);
    }
    if (wv == W_ANG) {
        const int nA = gsp.K - gsp.nR, a0 = gsp.a_poff[0];
        for (int q = lane; q < nA; q += 64) L.ci[q] = (unsigned char)P.a_cidx[a0 + q];
        if (FUSED)
            for (int k = lane; k < gsp.K; k += 64) {
                L.sc[k] = P.inpw64[gsp.in_off + k];
                L.sc[gsp.K + k] = P.inpb64[gsp.in_off + k];
            }
    }
    if (FUSED && wv > 0) {
        // the network parameters of the species, staged by waves 1.. (16-byte copies)
        const float4* src = (const float4*)(io.wpad + gsp.wp_off);
        float4* dst = (float4*)L.nw;
        const int n4 = gsp.wp_n >> 2;
        for (int e = tid - 64; e < n4; e += BT - 64) dst[e] = src[e];
    }

    // --- wave 0: decode the key row, displacements r_ij = posn(i) - (posn(j) + R n) in FP64
    // (nnp.cpp:363), drop pairs at or beyond the species' largest cutoff, bucket by species (ballot order
    // keeps the row order); slot q gets j in p[q] and (d, |d|) in u[q] ------------------------------------
    if (wv == 0) {
        const StrucInfo& S = io.sinfo[st];
        const double pix = io.pos[3 * i], piy = io.pos[3 * i + 1], piz = io.pos[3 * i + 2];
        const bool fits = nn <= nbm;
        // neighbour e of the row: atom j, species Jl (-1 past the cutoff or the row) and d, |d|
        auto decode = [&](int e, int& j, int& Jl, double& dx, double& dy, double& dz, double& r) {
            Jl = -1;
            j = 0;
            dx = dy = dz = r = 0.0;
            if (e < nn && fits) {
                const long long key = krow[e];
                const int jl = (int)(key >> 30), c = (int)(key & 0x3fffffff);
                const int c0 = (c >> 20) - 512, c1 = ((c >> 10) & 1023) - 512, c2 = (c & 1023) - 512;
                j = jl + S.atom_off;
                dx = pix - (io.pos[3 * j] + (S.R[0] * c0 + S.R[3] * c1 + S.R[6] * c2));
                dy = piy - (io.pos[3 * j + 1] + (S.R[1] * c0 + S.R[4] * c1 + S.R[7] * c2));
                dz = piz - (io.pos[3 * j + 2] + (S.R[2] * c0 + S.R[5] * c1 + S.R[8] * c2));
                r = sqrt(dx * dx + dy * dy + dz * dz);
                if ((float)r < rcut) Jl = io.type[j];
            }
        };
        int cnt[ACE_MAXT], fill[ACE_MAXT];
#pragma unroll
        for (int J = 0; J < ACE_MAXT; ++J) cnt[J] = 0;
        auto count = [&](int Jl) {
#pragma unroll
            for (int J = 0; J < ACE_MAXT; ++J)
                if (J < nt) cnt[J] += __popcll(__ballot(Jl == J));
        };
        // slot q of species Jl: the next free one in ballot order, i.e. the row order
        // the gathered gradient launches keep the row position e of slot q in p[q] instead of j, and store
        // zero gradients at the reverse entries of the row's pairs past the cutoff
        const bool pos_e = GACC && io.gin != nullptr;
        auto place = [&](int e, int j, int Jl, double dx, double dy, double dz, double r) {
            if (pos_e && Jl < 0 && e < nn) {
                const long long t = io.rev ? (long long)io.rev[(long long)i * io.cap + e]
                                           : reverse_entry(io.keys, io.cap, io.counts, krow[e], i - S.atom_off, S.atom_off);
                if (t >= 0) {
                    double* gp = io.gin + 3 * t;
                    gp[0] = 0.0;
                    gp[1] = 0.0;
                    gp[2] = 0.0;
                } else {
                    atomicAdd(io.err, 1);
                }
            }
#pragma unroll
            for (int J = 0; J < ACE_MAXT; ++J) {
                if (J < nt) {
                    const unsigned long long m = __ballot(Jl == J);
                    if (Jl == J) {
                        // unit vector, |d| and 1/|d| from the FP64 displacement, each rounded once
                        const int q = fill[J] + __popcll(m & lane_mask_lt());
                        const double ir = 1.0 / r;
                        if (LEAN) L.pe[q] = (unsigned short)e;   // e < cap <= 1024
                        else L.p[q] = pos_e ? e : j;
                        L.u[4 * q] = (T)(dx * ir);
                        L.u[4 * q + 1] = (T)(dy * ir);
                        L.u[4 * q + 2] = (T)(dz * ir);
                        L.u[4 * q + 3] = (T)ir;
                        if (LEAN) {
                            L.g[3 * nbm + q] = r;   // |d| in FP64 for the per-slot pass (wave 1's accumulators)
                        } else {
                            L.r[q] = (T)r;
                            if (!F64) L.rlo[q] = (float)(r - (double)(float)r);
                        }
                    }
                    fill[J] += __popcll(m);
                }
            }
        };
        // pass 0 counts per species; the first NCH chunks of 64 neighbours stay in registers for the
        // placing pass, so a row of up to 64 NCH neighbours is read once
        constexpr int NCH = 2;
        int jr[NCH], Jr[NCH];
        double dxr[NCH], dyr[NCH], dzr[NCH], rr[NCH];
#pragma unroll
        for (int c = 0; c < NCH; ++c) {
            decode(64 * c + lane, jr[c], Jr[c], dxr[c], dyr[c], dzr[c], rr[c]);
            count(Jr[c]);
        }
        for (int cb = 64 * NCH; cb < nn; cb += 64) {
            int j, Jl;
            double dx, dy, dz, r;
            decode(cb + lane, j, Jl, dx, dy, dz, r);
            count(Jl);
        }
        // pass 1: species offsets, then the slots
        {
            int acc = 0;
#pragma unroll
            for (int J = 0; J < ACE_MAXT; ++J) {
                fill[J] = acc;
                acc += (J < nt) ? cnt[J] : 0;
            }
            if (lane == 0) {
#pragma unroll
                for (int J = 0; J <= ACE_MAXT; ++J)
                    if (J <= nt) s_cnt[J] = (J < nt) ? fill[J] : acc;
            }
        }
        if (fits) {
#pragma unroll
            for (int c = 0; c < NCH; ++c) place(64 * c + lane, jr[c], Jr[c], dxr[c], dyr[c], dzr[c], rr[c]);
            for (int cb = 64 * NCH; cb < nn; cb += 64) {
                int j, Jl;
                double dx, dy, dz, r;
                decode(cb + lane, j, Jl, dx, dy, dz, r);
                place(cb + lane, j, Jl, dx, dy, dz, r);
            }
        }
        if (!fits && lane == 0) s_cnt[nt] = nbm + 1;
    }
    __syncthreads();
    const DevSpecies& sp = *(FUSED ? &s_sp : &gsp);
    const int ng = sp.ng;
    const int ntot = s_cnt[nt];
    if (ntot > nbm) {
        if (tid == 0) atomicAdd(io.err, 1);
        return;
    }

    // --- per-slot quantities ---------------------------------------------------------------------
    if (LEAN) {
        // cutoffs and angular factors, and the slot's radial dE/d|d| over the functions of its species block
        // (weights dE/dG in row, basis_radial.cpp:316-376) as wave 0's first gradient term
        for (int q = tid; q < ntot; q += BT) {
            int J = 0;
#pragma unroll
            for (int jj = 1; jj < ACE_MAXT; ++jj)
                if (jj < nt && q >= s_cnt[jj]) J = jj;
            const double r = L.g[3 * nbm + q];
            T fc, gc;
            t_cutoff_r64<T>(sp.r_cut[J], sp.r_rc64[J], r, fc, gc);
            T acc = T(0);
            const int nR = sp.r_n[J], goff = sp.r_goff[J], rq0 = sp.r_poff[J] - sp.r_poff[0], form = sp.r_form[J];
            if (form == ACE_RAD_GAUSSIAN)
                for (int n = 0; n < nR; ++n) {
                    T v, dv;
                    t_radial_r64<T>(ACE_RAD_GAUSSIAN, L.rp[rq0 + n], L.rp[sp.nR + rq0 + n], r, fc, gc, v, dv);
                    acc += (T)(ROWF ? L.rowf[goff + n] : L.row[goff + n]) * dv;
                }
            else
                for (int n = 0; n < nR; ++n) {
                    T v, dv;
                    t_radial_r64<T>(form, L.rp[rq0 + n], L.rp[sp.nR + rq0 + n], r, fc, gc, v, dv);
                    acc += (T)(ROWF ? L.rowf[goff + n] : L.row[goff + n]) * dv;
                }
            if (sp.a_cut != sp.r_cut[J] || sp.a_rc64 != sp.r_rc64[J]) t_cutoff_r64<T>(sp.a_cut, sp.a_rc64, r, fc, gc);
            const T fa = fc, ga = gc, rt = (T)r;
#pragma unroll
            for (int g = 0; g < NG; ++g) {
                T ev = T(0), edv = T(0);
                if (g < ng) {
                    const T ie = F64 ? (T)sp.ietar2_64[g] : (T)sp.ietar2[g];
                    const T e = t_exp<T>(-rt * rt * ie);
                    ev = fa * e;
                    edv = (ga - T(2) * rt * ie * fa) * e;
                }
                L.E[q * 2 * NG + g] = ev;
                L.E[q * 2 * NG + NG + g] = edv;
            }
            const double a = (double)acc;
            L.g[3 * q] = a * (double)L.u[4 * q];
            L.g[3 * q + 1] = a * (double)L.u[4 * q + 1];
            L.g[3 * q + 2] = a * (double)L.u[4 * q + 2];
        }
    } else
    for (int q = tid; q < ntot; q += BT) {
        int J = 0;
#pragma unroll
        for (int jj = 1; jj < ACE_MAXT; ++jj)
            if (jj < nt && q >= s_cnt[jj]) J = jj;
        // cutoffs from |d| to FP64 precision; the separable angular factors in T (their exponents are below
        // ~1 in magnitude)
        const double r = slot_r64(L, q);
        T fc, gc;
        t_cutoff_r64<T>(sp.r_cut[J], sp.r_rc64[J], r, fc, gc);
        L.fR[q] = fc;
        L.gR[q] = gc;
        if (sp.a_cut != sp.r_cut[J] || sp.a_rc64 != sp.r_rc64[J]) t_cutoff_r64<T>(sp.a_cut, sp.a_rc64, r, fc, gc);
        const T fa = fc, ga = gc, rt = L.r[q];
        if (GEN) {
            L.fA[q] = fa;
            L.gA[q] = ga;
        }
#pragma unroll
        for (int g = 0; g < NG; ++g) {
            T ev = T(0), edv = T(0);
            if (g < ng) {
                const T ie = F64 ? (T)sp.ietar2_64[g] : (T)sp.ietar2[g];
                const T e = t_exp<T>(-rt * rt * ie);
                ev = fa * e;
                edv = (ga - T(2) * rt * ie * fa) * e;
            }
            L.E[q * 2 * NG + g] = ev;
            L.E[q * 2 * NG + NG + g] = edv;
        }
        if (GACC) {
            L.fr[q] = T(0);
            for (int w = 0; w < BW; ++w) {
                L.g[w * 3 * nbm + 3 * q] = 0.0;
                L.g[w * 3 * nbm + 3 * q + 1] = 0.0;
                L.g[w * 3 * nbm + 3 * q + 2] = 0.0;
            }
        }
        if (TAN) {
            const int jj = L.p[q];
            const T ex = (T)(io.eps[3 * i] - io.eps[3 * jj]);
            const T ey = (T)(io.eps[3 * i + 1] - io.eps[3 * jj + 1]);
            const T ez = (T)(io.eps[3 * i + 2] - io.eps[3 * jj + 2]);
            L.te[4 * q] = ex;
            L.te[4 * q + 1] = ey;
            L.te[4 * q + 2] = ez;
            L.te[4 * q + 3] = L.u[4 * q] * ex + L.u[4 * q + 1] * ey + L.u[4 * q + 2] * ez;
        }
    }
    if (!ADJ)
        for (int k = tid; k < sp.K; k += BT) L.row[k] = 0.0;
    __syncthreads();

    if (FUSED) {
        basis_sums<T, NG, NZ, 0, GEN, BW, SPL>(L, sp, P, s_cnt, nt, nbm, tid, wv, lane);
        block_network<T>(L, sp, tid, &io.e_atom[i]);
        basis_sums<T, NG, NZ, 1, GEN, BW, SPL>(L, sp, P, s_cnt, nt, nbm, tid, wv, lane);
    } else {
        basis_sums<T, NG, NZ, MODE, GEN, BW, SPL, LEAN>(L, sp, P, s_cnt, nt, nbm, tid, wv, lane);
    }
    if (MODE == 0) {
        // network input row: the ANN's input scaling inpw*(G + inpb) (nn.cpp:1406) applied here, in FP64,
        // so the FP32 row keeps full relative precision after the mean is removed
        for (int k = tid; k < sp.K; k += BT) {
            const double w = P.inpw64[sp.in_off + k], b = P.inpb64[sp.in_off + k];
            io.G[(long long)row * io.ldg + k] = (float)(w * (L.row[k] + b));
        }
    } else if (TAN) {
        for (int k = tid; k < sp.K; k += BT) io.G[(long long)row * io.ldg + k] = (float)L.row[k];   // v
    } else {
        // g_q = dE_i/dd_q of slot q (pair i -> j): F_j -= -g_q, F_i -= g_q summed over the slots; the atom's
        // virial share -sum_q g_q (x) d_q; the atom's energy into its structure's total
        double fi[3] = {0.0, 0.0, 0.0}, w[9];
        const int a0i = io.sinfo[st].atom_off;
#pragma unroll
        for (int q9 = 0; q9 < 9; ++q9) w[q9] = 0.0;
        for (int q = tid; q < ntot; q += BT) {
            const double fr = LEAN ? 0.0 : (double)L.fr[q];   // the lean adjoint's radial part is in g
            const double ux = (double)L.u[4 * q], uy = (double)L.u[4 * q + 1], uz = (double)L.u[4 * q + 2];
            double g3[3] = {fr * ux, fr * uy, fr * uz};
            for (int ww = 0; ww < BW; ++ww) {
                g3[0] += (double)L.g[ww * 3 * nbm + 3 * q];
                g3[1] += (double)L.g[ww * 3 * nbm + 3 * q + 1];
                g3[2] += (double)L.g[ww * 3 * nbm + 3 * q + 2];
            }
            if (io.gin) {
                const int e = LEAN ? (int)L.pe[q] : L.p[q];
                const long long t = io.rev ? (long long)io.rev[(long long)i * io.cap + e]
                                           : reverse_entry(io.keys, io.cap, io.counts, krow[e], i - a0i, a0i);
                if (t >= 0) {
                    double* gp = io.gin + 3 * t;
                    gp[0] = g3[0];
                    gp[1] = g3[1];
                    gp[2] = g3[2];
                } else {
                    atomicAdd(io.err, 1);
                }
            } else {
                // the lean layout keeps the row position: j from the key
                const int j = LEAN ? (int)(krow[L.pe[q]] >> 30) + a0i : L.p[q];
                atomicAdd(&io.F_fix[3 * j], fix_of(g3[0]));
                atomicAdd(&io.F_fix[3 * j + 1], fix_of(g3[1]));
                atomicAdd(&io.F_fix[3 * j + 2], fix_of(g3[2]));
            }
            // |d| for the virial: the lean layout keeps 1/|d| only (its rounding is the unit vector's)
            const double r = LEAN ? 1.0 / (double)L.u[4 * q + 3] : slot_r64(L, q);
            const double d3[3] = {ux * r, uy * r, uz * r};
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                fi[a] -= g3[a];
#pragma unroll
                for (int b = 0; b < 3; ++b) w[3 * a + b] -= g3[a] * d3[b];
            }
        }
        double v[16];
#pragma unroll
        for (int a = 0; a < 3; ++a) v[a] = fi[a];
#pragma unroll
        for (int q9 = 0; q9 < 9; ++q9) v[3 + q9] = w[q9];
#pragma unroll
        for (int q4 = 12; q4 < 16; ++q4) v[q4] = 0.0;
        const double tot = wave_reduce_scatter<4, double>(v, lane);
        if ((lane & 3) == 0) L.sred[16 * wv + (lane >> 2)] = tot;
        __syncthreads();
        if (tid < 13) {
            double a = 0.0;
            if (tid < 12) {
                for (int ww = 0; ww < BW; ++ww) a += L.sred[16 * ww + tid];
            } else {
                a = io.e_atom[i];
            }
            if (tid < 3) {
                if (io.gin) io.fown[3 * i + tid] = a;
                else atomicAdd(&io.F_fix[3 * i + tid], fix_of(a));
            } else if (tid < 12) atomicAdd(&io.W_fix[9 * st + tid - 3], fix_of(a));
            else atomicAdd(&io.E_fix[st], fix_of(a));
        }
    }
}
