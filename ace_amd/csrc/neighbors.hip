// neighbors.hip -- periodic neighbour list on the GPU (cell list with periodic images).
//
// Produces exactly the set of NeighborList::build (src/struc/neighbor.cpp:46-87): every (j, n) with
// |(r_i - r_j) - R n|^2 < rc^2 evaluated in FP64 with the reference's formula and no contraction,
// j == i only when the distance exceeds 1e-12, n within ceil(3 rc / |a_d|) per dimension, and each
// atom's list in the reference order (j ascending, then n lexicographic over (-c..c)^3).
//
// Search: atoms are binned into cells of a periodic grid (fractional coordinates, nc_d cells along
// lattice vector d); the stencil half-width st_d covers rc along the perpendicular height, and a
// stencil cell past the grid edge stands for a translated copy of the cell (image shift), so small
// cells (L < 2 rc) yield every periodic image of a neighbour once.
//
// Pipeline (one build): k_nb_bin (workgroup per structure: cell counts, scan and a cell-ordered copy of
// the positions, all in LDS) -> k_nb_search (workgroup per atom; candidates are contiguous runs of the
// cell-ordered copy, loaded four chunks at a time; keys rank-sorted in LDS into the reference order and
// written to a fixed-capacity per-atom row, which the basis kernels read directly).  When every structure
// is small and its cell heights exceed twice the list radius, k_nb_brute replaces both (no cell list).
// CSR arrays are built from the rows on demand (export, unbatched port): exclusive scan of the counts ->
// k_nb_csr.
#include "ace_common.hpp"

// ---- binning: one workgroup per structure ---------------------------------------------------------

#define NB_BIN_THREADS 1024
#define NB_MAXC 8192   // cells per structure (the host coarsens the grid to stay within it)
#define NB_BIN_REG 4   // atoms per thread held in registers between the count and the scatter

// cell_start[cell_off + c] (global index into the cell-ordered arrays, per structure, with the end entry
// cell_start[cell_off + ncl] written by the structure's own workgroup at index cell_off + ncl; the
// host lays structures out with one spare entry each), cpos / cinfo: cell-ordered positions and
// (atom, wrap) pairs.  cell_of / wrap_of: per atom, in the atom's own slot.
__global__ __launch_bounds__(NB_BIN_THREADS) void k_nb_bin(const double* __restrict__ pos,
                                                           const StrucInfo* __restrict__ sinfo,
                                                           int* __restrict__ cell_of, int* __restrict__ wrap_of,
                                                           int* __restrict__ slot_of, int* __restrict__ cell_start,
                                                           double4* __restrict__ cpos, int2* __restrict__ cinfo,
                                                           int* __restrict__ flags) {
    __shared__ int cnt[NB_MAXC];
    __shared__ int wsum[NB_BIN_THREADS / 64];
    const StrucInfo& S = sinfo[blockIdx.x];
    const int t = threadIdx.x;
    const int ncl = S.nc[0] * S.nc[1] * S.nc[2];
    const int a0 = S.atom_off, na = S.natoms;
    const int cs0 = S.cell_off + blockIdx.x;          // this structure's cell_start block (ncl + 1 entries)
    if (blockIdx.x == 0 && t < 16) flags[t] = 0;      // build flags and counters (maximum, totals, errors)
    for (int c = t; c < ncl; c += NB_BIN_THREADS) cnt[c] = 0;
    __syncthreads();
    // the first NB_BIN_REG atoms of each thread keep position, cell, slot and wrap in registers for the
    // scatter below; atoms past them go through slot_of
    double rx[NB_BIN_REG], ry[NB_BIN_REG], rz[NB_BIN_REG];
    int rcell[NB_BIN_REG], rslot[NB_BIN_REG], rwrap[NB_BIN_REG];
    auto bin_one = [&](int k, double& x, double& y, double& z, int& cell, int& slot, int& wrap) {
        const int i = a0 + k;
        x = pos[3 * i];
        y = pos[3 * i + 1];
        z = pos[3 * i + 2];
        int c[3], w[3];
        for (int d = 0; d < 3; ++d) {
            const double f = S.Rinv[d] * x + S.Rinv[d + 3] * y + S.Rinv[d + 6] * z;
            const double fl = floor(f);
            int cc = (int)((f - fl) * S.nc[d]);
            cc = min(max(cc, 0), S.nc[d] - 1);
            c[d] = cc;
            w[d] = (int)fl;
        }
        cell = (c[2] * S.nc[1] + c[1]) * S.nc[0] + c[0];
        wrap = img_pack(w[0], w[1], w[2]);
        cell_of[i] = cell;
        wrap_of[i] = wrap;
        slot = atomicAdd(&cnt[cell], 1);
    };
#pragma unroll
    for (int u = 0; u < NB_BIN_REG; ++u) {
        const int k = t + u * NB_BIN_THREADS;
        if (k < na) bin_one(k, rx[u], ry[u], rz[u], rcell[u], rslot[u], rwrap[u]);
    }
    for (int k = t + NB_BIN_REG * NB_BIN_THREADS; k < na; k += NB_BIN_THREADS) {
        double x, y, z;
        int cell, slot, wrap;
        bin_one(k, x, y, z, cell, slot, wrap);
        slot_of[a0 + k] = slot;
    }
    __syncthreads();
    // exclusive scan of cnt: each thread sums a run of ceil(ncl / threads) cells, then a wave/block scan
    const int per = (ncl + NB_BIN_THREADS - 1) / NB_BIN_THREADS;
    const int c0 = t * per, c1 = min(c0 + per, ncl);
    int run = 0;
    for (int c = c0; c < c1; ++c) run += cnt[c];
    int incl = run;
    const int lane = t & 63, wv = t >> 6;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int v = __shfl_up(incl, o, 64);
        if (lane >= o) incl += v;
    }
    if (lane == 63) wsum[wv] = incl;
    __syncthreads();
    int wbase = 0;
    for (int q = 0; q < wv; ++q) wbase += wsum[q];
    int acc = a0 + wbase + incl - run;
    for (int c = c0; c < c1; ++c) {
        const int v = cnt[c];
        cnt[c] = acc;
        cell_start[cs0 + c] = acc;
        acc += v;
    }
    if (t == 0) cell_start[cs0 + ncl] = a0 + na;
    __syncthreads();
#pragma unroll
    for (int u = 0; u < NB_BIN_REG; ++u) {
        const int k = t + u * NB_BIN_THREADS;
        if (k < na) {
            const int dst = cnt[rcell[u]] + rslot[u];
            cpos[dst] = make_double4(rx[u], ry[u], rz[u], 0.0);
            cinfo[dst] = make_int2(a0 + k, rwrap[u]);
        }
    }
    for (int k = t + NB_BIN_REG * NB_BIN_THREADS; k < na; k += NB_BIN_THREADS) {
        const int i = a0 + k;
        const int dst = cnt[cell_of[i]] + slot_of[i];
        cpos[dst] = make_double4(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2], 0.0);
        cinfo[dst] = make_int2(i, wrap_of[i]);
    }
}

// ---- exclusive scan (int), three passes --------------------------------------------------------

#define SCAN_BLOCK 1024
__global__ void k_scan_local(int n, const int* __restrict__ in, int* __restrict__ out, int* __restrict__ block_sums) {
    __shared__ int s[SCAN_BLOCK];
    const int t = threadIdx.x;
    const int i = blockIdx.x * SCAN_BLOCK + t;
    const int v = (i < n) ? in[i] : 0;
    s[t] = v;
    __syncthreads();
    for (int off = 1; off < SCAN_BLOCK; off <<= 1) {
        const int a = (t >= off) ? s[t - off] : 0;
        __syncthreads();
        s[t] += a;
        __syncthreads();
    }
    if (i < n) out[i] = s[t] - v;
    if (t == SCAN_BLOCK - 1) block_sums[blockIdx.x] = s[t];
}

// single block: exclusive scan of up to SCAN_BLOCK*SCAN_BLOCK block sums is done in chunks
__global__ void k_scan_sums(int nb, int* __restrict__ block_sums, int* __restrict__ total, int n,
                            int* __restrict__ out) {
    __shared__ int s[SCAN_BLOCK];
    __shared__ int carry;
    const int t = threadIdx.x;
    if (t == 0) carry = 0;
    __syncthreads();
    for (int base = 0; base < nb; base += SCAN_BLOCK) {
        const int i = base + t;
        const int v = (i < nb) ? block_sums[i] : 0;
        s[t] = v;
        __syncthreads();
        for (int off = 1; off < SCAN_BLOCK; off <<= 1) {
            const int a = (t >= off) ? s[t - off] : 0;
            __syncthreads();
            s[t] += a;
            __syncthreads();
        }
        if (i < nb) block_sums[i] = s[t] - v + carry;
        __syncthreads();
        if (t == SCAN_BLOCK - 1) carry += s[t];
        __syncthreads();
    }
    if (t == 0) {
        *total = carry;
        out[n] = carry;
    }
}

__global__ void k_scan_add(int n, int* __restrict__ out, const int* __restrict__ block_sums) {
    const int i = blockIdx.x * SCAN_BLOCK + threadIdx.x;
    if (i < n) out[i] += block_sums[blockIdx.x];
}

// one workgroup: exclusive scan of n <= SCAN_BLOCK * SCAN_RUN ints; writes out[0..n] and *total
#define SCAN_RUN 64
__global__ __launch_bounds__(SCAN_BLOCK) void k_scan_single(int n, const int* __restrict__ in, int* __restrict__ out,
                                                            int* __restrict__ total) {
    __shared__ int wsum[SCAN_BLOCK / 64];
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    const int per = (n + SCAN_BLOCK - 1) / SCAN_BLOCK;
    const int c0 = min(t * per, n), c1 = min(c0 + per, n);
    int run = 0;
    for (int c = c0; c < c1; ++c) run += in[c];
    int incl = run;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int v = __shfl_up(incl, o, 64);
        if (lane >= o) incl += v;
    }
    if (lane == 63) wsum[wv] = incl;
    __syncthreads();
    int acc = incl - run;
    for (int q = 0; q < wv; ++q) acc += wsum[q];
    for (int c = c0; c < c1; ++c) {
        const int v = in[c];
        out[c] = acc;
        acc += v;
    }
    if (t == SCAN_BLOCK - 1) {
        out[n] = acc;
        *total = acc;
    }
}

// ---- search ---------------------------------------------------------------------------------------

__device__ inline int floordiv(int a, int b) { return (a >= 0) ? a / b : -((-a + b - 1) / b); }

// reference distance test (neighbor.cpp:72-82), FP64 without contraction: the lattice shift R n, then the
// squared distance of (dx, dy, dz) - R n
__device__ inline void ref_shift(const double* R, int n0, int n1, int n2, double& s0, double& s1, double& s2) {
#pragma clang fp contract(off)
    const double f0 = 1.0 * n0, f1 = 1.0 * n1, f2 = 1.0 * n2;
    s0 = R[0] * f0 + R[3] * f1 + R[6] * f2;
    s1 = R[1] * f0 + R[4] * f1 + R[7] * f2;
    s2 = R[2] * f0 + R[5] * f1 + R[8] * f2;
}
__device__ inline double ref_ds2_shifted(double dx, double dy, double dz, double s0, double s1, double s2) {
#pragma clang fp contract(off)
    const double v0 = dx - s0, v1 = dy - s1, v2 = dz - s2;
    return (v0 * v0 + v1 * v1) + v2 * v2;
}
__device__ inline double ref_ds2(const double* R, double dx, double dy, double dz, int n0, int n1, int n2) {
    double s0, s1, s2;
    ref_shift(R, n0, n1, n2, s0, s1, s2);
    return ref_ds2_shifted(dx, dy, dz, s0, s1, s2);
}

// neighbour order key: j_local in the high bits, then the image (n0, n1, n2) offset by 512 in three
// 10-bit fields; integer order = the reference order (j ascending, then n lexicographic)
__device__ inline long long nb_key(int jl, int n0, int n1, int n2) {
    return ((long long)jl << 30) | ((long long)(n0 + 512) << 20) | ((n1 + 512) << 10) | (n2 + 512);
}
__device__ inline void nb_key_decode(long long k, int& jl, int& n0, int& n1, int& n2) {
    jl = (int)(k >> 30);
    const int c = (int)(k & 0x3fffffff);
    n0 = (c >> 20) - 512;
    n1 = ((c >> 10) & 1023) - 512;
    n2 = (c & 1023) - 512;
}

// Rank sort of one atom's keys with order-preserving 32-bit keys: (j_local, n0 + m, n1 + m, n2 + m) packed
// in fields of b bits for the offsets (2 m < 2^b, m the largest image range of the structure) and the rest
// for j_local.  Thread t holds keys t + k BT (k < NK): it reads them, the workgroup overwrites the front of
// the key buffer with the 32-bit keys (padded to a multiple of four with 0xffffffff, larger than any key:
// a key of all ones needs n + m = 2^b - 1, odd, above 2 m), and every thread counts the smaller ones for
// its NK keys in one sweep.  count <= NK BT, and count rounded up to four stays within the buffer (cap is
// a multiple of eight).
template <int NK, int BT>
__device__ inline void nb_rank_sort32(long long* s_key, int count, int tid, int m, int b, long long* row) {
    long long a[NK];
    unsigned c[NK];
#pragma unroll
    for (int k = 0; k < NK; ++k) {
        const int e = tid + k * BT;
        a[k] = 0;
        c[k] = 0xffffffffu;
        if (e < count) {
            a[k] = s_key[e];
            int jl, n0, n1, n2;
            nb_key_decode(a[k], jl, n0, n1, n2);
            c[k] = ((unsigned)jl << (3 * b)) | ((unsigned)(n0 + m) << (2 * b)) | ((unsigned)(n1 + m) << b) |
                   (unsigned)(n2 + m);
        }
    }
    __syncthreads();
    unsigned* s_c = reinterpret_cast<unsigned*>(s_key);
    const int count4 = (count + 3) & ~3;
#pragma unroll
    for (int k = 0; k < NK; ++k)
        if (tid + k * BT < count4) s_c[tid + k * BT] = c[k];
    __syncthreads();
    int r[NK];
#pragma unroll
    for (int k = 0; k < NK; ++k) r[k] = 0;
    const uint2* s_c2 = reinterpret_cast<const uint2*>(s_key);
    for (int f = 0; f < count4; f += 4) {
        const uint2 p = s_c2[f >> 1], q = s_c2[(f >> 1) + 1];
#pragma unroll
        for (int k = 0; k < NK; ++k)
            r[k] += (int)(p.x < c[k]) + (int)(p.y < c[k]) + (int)(q.x < c[k]) + (int)(q.y < c[k]);
    }
#pragma unroll
    for (int k = 0; k < NK; ++k)
        if (tid + k * BT < count) __builtin_nontemporal_store(a[k], &row[r[k]]);   // read once, by the next kernel
}

#define NB_WAVES 4          // waves per atom of the binning-free search
#define NB_SEARCH_WAVES 1   // waves per atom of the binned search
#define NB_LMAX 1024   // neighbours per atom held while sorting
#define NB_U 2         // candidate chunks loaded together by one wave
#define NB_SC 128      // stencil cells per pass (two per lane): the 5 x 5 x 5 stencil in one pass
#define NB_MAP 1024    // candidates of a pass mapped to their stencil cell in LDS (longer passes search the prefix)

// Workgroup of NB_SEARCH_WAVES waves per atom (one: the most atoms in flight per CU, which hides the dependent
// loads of the stencil and the candidates best); the stencil is taken NB_SC cells at a time (each pass is two
// dependent loads: cell table, then candidates), 64-candidate chunks are dealt to the waves in turn, NB_U chunks
// in flight per wave.
// MODE 0: count only (counts[i], running maximum in *maxcount).  MODE 1: sort the keys
// (j_local << 30 | image order code: the reference order) and write them to row i of keys (cap slots); an
// atom with more than cap neighbours raises *over and writes none.
template <int MODE>
// 8 waves per SIMD (its SGPRs alone would allow 7; a few spill to VGPR lanes)
__global__ __launch_bounds__(NB_SEARCH_WAVES * 64) __attribute__((amdgpu_waves_per_eu(8))) void k_nb_search(
    int natoms, double rc2, const double* __restrict__ pos, const int* __restrict__ struc_of,
    const StrucInfo* __restrict__ sinfo, const int* __restrict__ cell_of, const int* __restrict__ wrap_of,
    const int* __restrict__ cell_start, const double4* __restrict__ cpos, const int2* __restrict__ cinfo,
    int* __restrict__ counts, int* __restrict__ maxcount, long long* __restrict__ keys, int cap,
    int* __restrict__ over) {
    __shared__ int s_cstart[NB_SC];
    __shared__ int s_cpref[NB_SC + 1];
    __shared__ int s_cimg[NB_SC];
    __shared__ __align__(8) unsigned char s_slot[NB_MAP];   // after the search: the j_local bitmap
    __shared__ double4 s_rn[27];           // R n for n in {-1, 0, 1}^3, index (n0 + 1) + 3 (n1 + 1) + 9 (n2 + 1)
    __shared__ int s_n;
    extern __shared__ long long s_key[];   // MODE 1: cap keys (dynamic LDS)
    const int tid = threadIdx.x, wv = tid >> 6, lane = tid & 63;
    const int i = blockIdx.x;
    const int s = struc_of[i];
    const StrucInfo& S = sinfo[s];
    const int cs0 = S.cell_off + s;
    const int loc = cell_of[i];
    const int ci[3] = {loc % S.nc[0], (loc / S.nc[0]) % S.nc[1], loc / (S.nc[0] * S.nc[1])};
    int wi0, wi1, wi2;
    img_unpack(wrap_of[i], wi0, wi1, wi2);
    const double pix = pos[3 * i], piy = pos[3 * i + 1], piz = pos[3 * i + 2];
    const int sx = 2 * S.st[0] + 1, sy = 2 * S.st[1] + 1, sz = 2 * S.st[2] + 1;
    const int nst = sx * sy * sz;
    // stencils within one image of the cell grid (st <= nc): the cell of stencil slot q by multiply-high with
    // m = floor((2^32 - 1) / d) + 1 (exact while q e < 2^32, e = m d - 2^32 < d) and one-image wrapping
    const bool fast = S.st[0] <= S.nc[0] && S.st[1] <= S.nc[1] && S.st[2] <= S.nc[2] && nst < (1 << 20) &&
                      sx * sy < 4096;
    const unsigned mx = 0xffffffffu / (unsigned)sx + 1u, mxy = 0xffffffffu / (unsigned)(sx * sy) + 1u;
    const int cf0 = S.cf[0], cf1 = S.cf[1], cf2 = S.cf[2], aoff = S.atom_off;
    if (tid == 0) s_n = 0;
    if (tid < 27) {
        double s0, s1, s2;
        ref_shift(S.R, tid % 3 - 1, (tid / 3) % 3 - 1, tid / 9 - 1, s0, s1, s2);
        s_rn[tid] = make_double4(s0, s1, s2, 0.0);
    }
    for (int cb = 0; cb < nst; cb += NB_SC) {
        // stencil cells cb + 2 lane and cb + 2 lane + 1 of this pass: start, size and image shift, with the
        // exclusive prefix of the candidate counts (wave 0)
        __syncthreads();
        if (wv == 0) {
            int cnt[2], start[2], shift[2];
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int q = cb + 2 * lane + h;
                cnt[h] = 0;
                start[h] = 0;
                shift[h] = 0;
                if (q < nst) {
                    int a0, a1, a2, n0, n1, n2;
                    if (fast) {
                        const int qy = (int)__umulhi((unsigned)q, mx), qz = (int)__umulhi((unsigned)q, mxy);
                        a0 = ci[0] + q - qy * sx - S.st[0];
                        a1 = ci[1] + qy - qz * sy - S.st[1];
                        a2 = ci[2] + qz - S.st[2];
                        n0 = a0 < 0 ? -1 : (a0 >= S.nc[0] ? 1 : 0);
                        n1 = a1 < 0 ? -1 : (a1 >= S.nc[1] ? 1 : 0);
                        n2 = a2 < 0 ? -1 : (a2 >= S.nc[2] ? 1 : 0);
                    } else {
                        a0 = ci[0] + q % sx - S.st[0];
                        a1 = ci[1] + (q / sx) % sy - S.st[1];
                        a2 = ci[2] + q / (sx * sy) - S.st[2];
                        n0 = floordiv(a0, S.nc[0]);
                        n1 = floordiv(a1, S.nc[1]);
                        n2 = floordiv(a2, S.nc[2]);
                    }
                    const int c0 = a0 - n0 * S.nc[0], c1 = a1 - n1 * S.nc[1], c2 = a2 - n2 * S.nc[2];
                    const int cell = cs0 + (c2 * S.nc[1] + c1) * S.nc[0] + c0;
                    start[h] = cell_start[cell];
                    cnt[h] = cell_start[cell + 1] - start[h];
                    shift[h] = img_pack(n0, n1, n2);
                }
            }
            const int two = cnt[0] + cnt[1];
            int pre = two;
#pragma unroll
            for (int o = 1; o < 64; o <<= 1) {
                const int t = __shfl_up(pre, o, 64);
                if (lane >= o) pre += t;
            }
            s_cstart[2 * lane] = start[0];
            s_cstart[2 * lane + 1] = start[1];
            s_cimg[2 * lane] = shift[0];
            s_cimg[2 * lane + 1] = shift[1];
            s_cpref[2 * lane] = pre - two;
            s_cpref[2 * lane + 1] = pre - two + cnt[0];
            if (lane == 63) s_cpref[NB_SC] = pre;
            // stencil slot of every candidate of the pass, when they fit the map
            if (__shfl(pre, 63, 64) <= NB_MAP) {
                for (int t = 0; t < cnt[0]; ++t) s_slot[pre - two + t] = (unsigned char)(2 * lane);
                for (int t = 0; t < cnt[1]; ++t) s_slot[pre - cnt[1] + t] = (unsigned char)(2 * lane + 1);
            }
        }
        __syncthreads();
        const int total = s_cpref[NB_SC];
        const bool mapped = total <= NB_MAP;
        // chunk q of 64 candidates goes to wave q mod NB_SEARCH_WAVES, so a short stencil keeps every wave busy
        for (int tb = 0; tb < total; tb += NB_SEARCH_WAVES * 64 * NB_U) {
            double4 pj[NB_U];
            int2 inf[NB_U];
            int sl[NB_U];
#pragma unroll
            for (int u = 0; u < NB_U; ++u) {
                const int x = tb + 64 * (u * NB_SEARCH_WAVES + wv) + lane;
                // stencil slot of candidate x: largest c with s_cpref[c] <= x
                int c = 0;
                if (mapped) {
                    if (x < total) c = s_slot[x];
                } else {
#pragma unroll
                    for (int st = NB_SC / 2; st > 0; st >>= 1)
                        if (s_cpref[c + st] <= x) c += st;
                }
                sl[u] = c;
                if (x < total) {
                    const int e = s_cstart[c] + (x - s_cpref[c]);
                    pj[u] = cpos[e];
                    inf[u] = cinfo[e];
                }
            }
#pragma unroll
            for (int u = 0; u < NB_U; ++u) {
                const int x = tb + 64 * (u * NB_SEARCH_WAVES + wv) + lane;
                bool acc = false;
                long long key = 0;
                if (x < total) {
                    const int j = inf[u].x;
                    int sh0, sh1, sh2, wj0, wj1, wj2;
                    img_unpack(s_cimg[sl[u]], sh0, sh1, sh2);
                    img_unpack(inf[u].y, wj0, wj1, wj2);
                    const int n0 = sh0 + wi0 - wj0, n1 = sh1 + wi1 - wj1, n2 = sh2 + wi2 - wj2;
                    // |n_d| <= cf_d, one unsigned comparison per dimension
                    if (((unsigned)(n0 + cf0) <= (unsigned)(2 * cf0)) & ((unsigned)(n1 + cf1) <= (unsigned)(2 * cf1)) &
                        ((unsigned)(n2 + cf2) <= (unsigned)(2 * cf2))) {
                        double dx, dy, dz;
                        {
#pragma clang fp contract(off)
                            dx = pix - pj[u].x;
                            dy = piy - pj[u].y;
                            dz = piz - pj[u].z;
                        }
                        // R n from the table for the one-image shifts (the same operations, so the same bits)
                        double s0, s1, s2;
                        const unsigned t0 = n0 + 1, t1 = n1 + 1, t2 = n2 + 1;
                        if ((t0 <= 2u) & (t1 <= 2u) & (t2 <= 2u)) {
                            const double4 r = s_rn[t0 + 3 * t1 + 9 * t2];
                            s0 = r.x;
                            s1 = r.y;
                            s2 = r.z;
                        } else {
                            ref_shift(S.R, n0, n1, n2, s0, s1, s2);
                        }
                        const double ds2 = ref_ds2_shifted(dx, dy, dz, s0, s1, s2);
                        if ((j != i || ds2 > 1e-12) && ds2 < rc2) {
                            acc = true;
                            key = nb_key(j - aoff, n0, n1, n2);
                        }
                    }
                }
                const unsigned long long m = __ballot(acc);
                int base = 0;
                if (lane == 0 && m) base = atomicAdd(&s_n, __popcll(m));
                base = __shfl(base, 0, 64);
                if (MODE == 1 && acc) {
                    const int r = base + __popcll(m & lane_mask_lt());
                    if (r < cap) s_key[r] = key;
                }
            }
        }
    }
    __syncthreads();
    const int count = s_n;
    if (tid == 0) {
        counts[i] = count;
        // the running maximum is only read after the counting pass; one address updated by every
        // workgroup serialises in L2 and delays the end of the kernel, so the filling pass skips it
        if (MODE == 0) atomicMax(maxcount, count);
    }
    if (MODE == 0) return;
    if (count > cap || count > NB_LMAX) {
        if (tid == 0) atomicMax(over, 1);
        return;
    }
    long long* row = keys + (long long)i * cap;
    constexpr int BT = NB_SEARCH_WAVES * 64;
    // Rows whose neighbours are distinct atoms (one image each: every cell height over twice the radius does
    // it) of a structure of at most 8 NB_MAP atoms: the rank of key (j, n) is the number of neighbours with a
    // smaller j, read from a bitmap of j_local (in the map's LDS, free after the search) and an exclusive
    // prefix of its words' popcounts (in the prefix's LDS).  A repeated j sends the row to the rank sort.
    if (S.natoms <= 8 * NB_MAP) {
        unsigned long long* bits = reinterpret_cast<unsigned long long*>(s_slot);
        int* wpre = s_cpref;
        const int nw = (S.natoms + 63) >> 6;
        __syncthreads();
        for (int w = tid; w < nw; w += BT) bits[w] = 0ull;
        if (tid == 0) s_n = 0;
        __syncthreads();
        for (int e = tid; e < count; e += BT) {
            const int jl = (int)(s_key[e] >> 30);
            const unsigned long long bit = 1ull << (jl & 63);
            if (atomicOr(&bits[jl >> 6], bit) & bit) s_n = 1;   // benign race: every writer stores 1
        }
        __syncthreads();
        if (s_n == 0) {
            if (wv == 0) {
                // exclusive prefix of the popcounts, two words per lane (nw <= 2 * 64)
                const int c0 = 2 * lane < nw ? __popcll(bits[2 * lane]) : 0;
                const int c1 = 2 * lane + 1 < nw ? __popcll(bits[2 * lane + 1]) : 0;
                int pre = c0 + c1;
#pragma unroll
                for (int o = 1; o < 64; o <<= 1) {
                    const int t = __shfl_up(pre, o, 64);
                    if (lane >= o) pre += t;
                }
                wpre[2 * lane] = pre - c0 - c1;
                wpre[2 * lane + 1] = pre - c1;
            }
            __syncthreads();
            for (int e = tid; e < count; e += BT) {
                const long long k = s_key[e];
                const int jl = (int)(k >> 30), w = jl >> 6;
                const int r = wpre[w] + __popcll(bits[w] & ((1ull << (jl & 63)) - 1ull));
                __builtin_nontemporal_store(k, &row[r]);   // read once, by the next kernel
            }
            return;
        }
    }
    // rank sort into the reference order (the keys of one atom are distinct): with 32-bit keys when they
    // hold j_local and the image offsets and the row fits four keys per thread, else with the full keys
    const int m = max(S.cf[0], max(S.cf[1], S.cf[2]));
    const int b = 32 - __clz(2 * m);                  // bits of 0 .. 2 m
    const int nk = (count + BT - 1) / BT;
    if (3 * b < 32 && (long long)S.natoms <= (1LL << (32 - 3 * b)) && nk <= 4) {
        if (nk <= 1) nb_rank_sort32<1, BT>(s_key, count, tid, m, b, row);
        else if (nk == 2) nb_rank_sort32<2, BT>(s_key, count, tid, m, b, row);
        else if (nk == 3) nb_rank_sort32<3, BT>(s_key, count, tid, m, b, row);
        else nb_rank_sort32<4, BT>(s_key, count, tid, m, b, row);
        return;
    }
    for (int e = tid; e < count; e += BT) {
        const long long k = s_key[e];
        int rank = 0;
        for (int f = 0; f < count; ++f) rank += (s_key[f] < k);
        __builtin_nontemporal_store(k, &row[rank]);   // read once, by the next kernel
    }
}

// Small periodic structures whose perpendicular heights all exceed twice the list radius: every atom of
// the structure is a candidate, under the minimum image, and no binning kernel runs.  At most one image
// of j lies within the radius (two images differ by a lattice vector longer than twice it), and each
// component of its fractional offset from r_i - r_j is below 1/2, so rounding the fractional difference
// finds it; the image range, the distance test and the key order are those of k_nb_search.  The build's
// list-overflow flag is *over; the kernel clears *over_next (the flag of the next build) and the basis
// kernels' refusal flag, which the binning kernel clears otherwise.
#define NB_BRUTE_MAX 4096   // atoms per structure for the brute-force search
template <int MODE>
__global__ __launch_bounds__(NB_WAVES * 64) void k_nb_brute(
    int natoms, double rc2, const double* __restrict__ pos, const int* __restrict__ struc_of,
    const StrucInfo* __restrict__ sinfo, int* __restrict__ counts, int* __restrict__ maxcount,
    long long* __restrict__ keys, int cap, int* __restrict__ over, int* __restrict__ over_next,
    int* __restrict__ basis_err) {
    __shared__ int s_n;
    extern __shared__ long long s_key[];   // MODE 1: cap keys (dynamic LDS)
    const int tid = threadIdx.x, lane = tid & 63;
    const int i = blockIdx.x;
    if (i == 0 && tid == 0) {
        *over_next = 0;
        *basis_err = 0;
    }
    const StrucInfo& S = sinfo[struc_of[i]];
    const int a0 = S.atom_off, na = S.natoms;
    const double pix = pos[3 * i], piy = pos[3 * i + 1], piz = pos[3 * i + 2];
    if (tid == 0) s_n = 0;
    __syncthreads();
    for (int cb = 0; cb < na; cb += NB_WAVES * 64) {
        const int x = cb + tid;   // candidate j = a0 + x
        bool acc = false;
        long long key = 0;
        if (x < na) {
            const int j = a0 + x;
            double dx, dy, dz;
            {
#pragma clang fp contract(off)
                dx = pix - pos[3 * j];
                dy = piy - pos[3 * j + 1];
                dz = piz - pos[3 * j + 2];
            }
            const int n0 = (int)rint(S.Rinv[0] * dx + S.Rinv[3] * dy + S.Rinv[6] * dz);
            const int n1 = (int)rint(S.Rinv[1] * dx + S.Rinv[4] * dy + S.Rinv[7] * dz);
            const int n2 = (int)rint(S.Rinv[2] * dx + S.Rinv[5] * dy + S.Rinv[8] * dz);
            if (n0 >= -S.cf[0] && n0 <= S.cf[0] && n1 >= -S.cf[1] && n1 <= S.cf[1] && n2 >= -S.cf[2] &&
                n2 <= S.cf[2]) {
                const double ds2 = ref_ds2(S.R, dx, dy, dz, n0, n1, n2);
                if ((j != i || ds2 > 1e-12) && ds2 < rc2) {
                    acc = true;
                    key = nb_key(x, n0, n1, n2);
                }
            }
        }
        const unsigned long long m = __ballot(acc);
        int base = 0;
        if (lane == 0 && m) base = atomicAdd(&s_n, __popcll(m));
        base = __shfl(base, 0, 64);
        if (MODE == 1 && acc) {
            const int r = base + __popcll(m & lane_mask_lt());
            if (r < cap) s_key[r] = key;
        }
    }
    __syncthreads();
    const int count = s_n;
    if (tid == 0) {
        counts[i] = count;
        if (MODE == 0) atomicMax(maxcount, count);
    }
    if (MODE == 0) return;
    if (count > cap || count > NB_LMAX) {
        if (tid == 0) atomicMax(over, 1);
        return;
    }
    // At most one image of each neighbour lies within the radius here, so the rank of key (j, n) in the
    // reference order is the number of neighbours with a smaller j: set bits below j in a bitmap of j_local
    // plus the exclusive prefix of the words' popcounts.
    __shared__ unsigned long long s_bits[NB_BRUTE_MAX / 64];
    __shared__ int s_wpre[NB_BRUTE_MAX / 64];
    constexpr int BT = NB_WAVES * 64;
    long long* row = keys + (long long)i * cap;
    const int nw = (na + 63) >> 6;
    for (int w = tid; w < nw; w += BT) s_bits[w] = 0ull;
    __syncthreads();
    for (int e = tid; e < count; e += BT) {
        const int jl = (int)(s_key[e] >> 30);
        atomicOr(&s_bits[jl >> 6], 1ull << (jl & 63));
    }
    __syncthreads();
    if (tid < 64) {
        const int c = tid < nw ? __popcll(s_bits[tid]) : 0;
        int pre = c;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const int t = __shfl_up(pre, o, 64);
            if (lane >= o) pre += t;
        }
        s_wpre[tid] = pre - c;
    }
    __syncthreads();
    for (int e = tid; e < count; e += BT) {
        const long long k = s_key[e];
        const int jl = (int)(k >> 30), w = jl >> 6;
        __builtin_nontemporal_store(k, &row[s_wpre[w] + __popcll(s_bits[w] & ((1ull << (jl & 63)) - 1ull))]);
    }
}

// Wave per atom: CSR arrays (neighbour index and packed image, the layout of NeighborList's accessors,
// neighbor.hpp:34-40) from the sorted key rows, for export and for the unbatched port.
__global__ __launch_bounds__(256) void k_nb_csr(int natoms, const int* __restrict__ struc_of,
                                                const StrucInfo* __restrict__ sinfo, const long long* __restrict__ keys,
                                                int cap, const int* __restrict__ counts, const int* __restrict__ nb_off,
                                                int* __restrict__ nb_j, int* __restrict__ nb_img) {
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int i = blockIdx.x * 4 + wv;
    if (i >= natoms) return;
    const int a0 = sinfo[struc_of[i]].atom_off;
    const int base = nb_off[i], n = counts[i];
    const long long* row = keys + (long long)i * cap;
    for (int e = lane; e < n; e += 64) {
        int jl, c0, c1, c2;
        nb_key_decode(row[e], jl, c0, c1, c2);
        nb_j[base + e] = jl + a0;
        nb_img[base + e] = img_pack(c0, c1, c2);
    }
}
