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
#include "ace_common.hpp"

// ---- binning -------------------------------------------------------------------------------------

// thread per atom: fractional coordinates, wrap vector, cell index
__global__ void k_nb_assign(int natoms, const double* __restrict__ pos, const int* __restrict__ struc_of,
                            const StrucInfo* __restrict__ sinfo, int* __restrict__ cell_of,
                            int* __restrict__ wrap_of, int* __restrict__ cell_count) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= natoms) return;
    const StrucInfo& S = sinfo[struc_of[i]];
    const double x = pos[3 * i], y = pos[3 * i + 1], z = pos[3 * i + 2];
    int c[3], w[3];
    for (int d = 0; d < 3; ++d) {
        const double f = S.Rinv[d] * x + S.Rinv[d + 3] * y + S.Rinv[d + 6] * z;
        const double fl = floor(f);
        double fw = f - fl;
        int cc = (int)(fw * S.nc[d]);
        if (cc >= S.nc[d]) cc = S.nc[d] - 1;
        if (cc < 0) cc = 0;
        c[d] = cc;
        w[d] = (int)fl;
    }
    const int cell = S.cell_off + (c[2] * S.nc[1] + c[1]) * S.nc[0] + c[0];
    cell_of[i] = cell;
    wrap_of[i] = img_pack(w[0], w[1], w[2]);
    atomicAdd(&cell_count[cell], 1);
}

__global__ void k_nb_scatter(int natoms, const int* __restrict__ cell_of, const int* __restrict__ cell_start,
                             int* __restrict__ cell_cursor, int* __restrict__ cell_atoms) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= natoms) return;
    const int c = cell_of[i];
    const int slot = atomicAdd(&cell_cursor[c], 1);
    cell_atoms[cell_start[c] + slot] = i;
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

// ---- search ---------------------------------------------------------------------------------------

__device__ inline int floordiv(int a, int b) { return (a >= 0) ? a / b : -((-a + b - 1) / b); }

// reference distance test (neighbor.cpp:72-82), FP64 without contraction
__device__ inline double ref_ds2(const double* R, double dx, double dy, double dz, int n0, int n1, int n2) {
#pragma clang fp contract(off)
    const double f0 = 1.0 * n0, f1 = 1.0 * n1, f2 = 1.0 * n2;
    const double s0 = R[0] * f0 + R[3] * f1 + R[6] * f2;
    const double s1 = R[1] * f0 + R[4] * f1 + R[7] * f2;
    const double s2 = R[2] * f0 + R[5] * f1 + R[8] * f2;
    const double v0 = dx - s0, v1 = dy - s1, v2 = dz - s2;
    return (v0 * v0 + v1 * v1) + v2 * v2;
}

#define NB_WAVES 4
#define NB_LMAX 1024   // neighbours per atom held while sorting

// Wave per atom.  mode 0: count (writes counts[i], max into *maxcount); mode 1: fill, sorted.
template <int MODE>
__global__ __launch_bounds__(NB_WAVES * 64) void k_nb_search(
    int natoms, double rc2, const double* __restrict__ pos, const int* __restrict__ struc_of,
    const StrucInfo* __restrict__ sinfo, const int* __restrict__ cell_of, const int* __restrict__ wrap_of,
    const int* __restrict__ cell_start, const int* __restrict__ cell_atoms, int* __restrict__ counts,
    int* __restrict__ maxcount, const int* __restrict__ nb_off, int* __restrict__ nb_j, int* __restrict__ nb_img,
    float4* __restrict__ nb_d, int* __restrict__ nb_src) {
    __shared__ int s_cstart[NB_WAVES][64];
    __shared__ int s_cpref[NB_WAVES][65];
    __shared__ int s_cimg[NB_WAVES][64];
    __shared__ long long s_key[MODE ? NB_WAVES : 1][MODE ? NB_LMAX : 1];
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int i = blockIdx.x * NB_WAVES + wv;
    if (i >= natoms) return;
    const int s = struc_of[i];
    const StrucInfo& S = sinfo[s];
    const int loc = cell_of[i] - S.cell_off;
    const int ci[3] = {loc % S.nc[0], (loc / S.nc[0]) % S.nc[1], loc / (S.nc[0] * S.nc[1])};
    int wi0, wi1, wi2;
    img_unpack(wrap_of[i], wi0, wi1, wi2);
    const double pix = pos[3 * i], piy = pos[3 * i + 1], piz = pos[3 * i + 2];
    const int sx = 2 * S.st[0] + 1, sy = 2 * S.st[1] + 1, sz = 2 * S.st[2] + 1;
    const int nst = sx * sy * sz;
    const int ny = 2 * S.cf[1] + 1, nz = 2 * S.cf[2] + 1;
    const long long nimg = (long long)(2 * S.cf[0] + 1) * ny * nz;
    int count = 0;
    for (int cb = 0; cb < nst; cb += 64) {
        // stencil cells of this chunk
        const int q = cb + lane;
        int cnt = 0, start = 0, shift = 0;
        if (q < nst) {
            const int d0 = q % sx - S.st[0], d1 = (q / sx) % sy - S.st[1], d2 = q / (sx * sy) - S.st[2];
            const int a0 = ci[0] + d0, a1 = ci[1] + d1, a2 = ci[2] + d2;
            const int n0 = floordiv(a0, S.nc[0]), n1 = floordiv(a1, S.nc[1]), n2 = floordiv(a2, S.nc[2]);
            const int c0 = a0 - n0 * S.nc[0], c1 = a1 - n1 * S.nc[1], c2 = a2 - n2 * S.nc[2];
            const int cell = S.cell_off + (c2 * S.nc[1] + c1) * S.nc[0] + c0;
            start = cell_start[cell];
            cnt = cell_start[cell + 1] - start;
            shift = img_pack(n0, n1, n2);
        }
        // inclusive prefix over lanes
        int pre = cnt;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const int t = __shfl_up(pre, o, 64);
            if (lane >= o) pre += t;
        }
        s_cstart[wv][lane] = start;
        s_cimg[wv][lane] = shift;
        s_cpref[wv][lane + 1] = pre;
        if (lane == 0) s_cpref[wv][0] = 0;
        const int total = __shfl(pre, 63, 64);
        wave_sync();
        for (int tb = 0; tb < total; tb += 64) {
            const int x = tb + lane;
            bool acc = false;
            long long key = 0;
            if (x < total) {
                // cell slot: largest k with pref[k] <= x
                int lo = 0, hi = 63;
                while (lo < hi) {
                    const int mid = (lo + hi + 1) >> 1;
                    if (s_cpref[wv][mid] <= x) lo = mid; else hi = mid - 1;
                }
                const int j = cell_atoms[s_cstart[wv][lo] + (x - s_cpref[wv][lo])];
                int sh0, sh1, sh2, wj0, wj1, wj2;
                img_unpack(s_cimg[wv][lo], sh0, sh1, sh2);
                img_unpack(wrap_of[j], wj0, wj1, wj2);
                const int n0 = sh0 + wi0 - wj0, n1 = sh1 + wi1 - wj1, n2 = sh2 + wi2 - wj2;
                if (n0 >= -S.cf[0] && n0 <= S.cf[0] && n1 >= -S.cf[1] && n1 <= S.cf[1] &&
                    n2 >= -S.cf[2] && n2 <= S.cf[2]) {
                    double dx, dy, dz;
                    {
#pragma clang fp contract(off)
                        dx = pix - pos[3 * j];
                        dy = piy - pos[3 * j + 1];
                        dz = piz - pos[3 * j + 2];
                    }
                    const double ds2 = ref_ds2(S.R, dx, dy, dz, n0, n1, n2);
                    if ((j != i || ds2 > 1e-12) && ds2 < rc2) {
                        acc = true;
                        const long long code = ((long long)(n0 + S.cf[0]) * ny + (n1 + S.cf[1])) * nz + (n2 + S.cf[2]);
                        key = (long long)(j - S.atom_off) * nimg + code;
                    }
                }
            }
            const unsigned long long m = __ballot(acc);
            if (MODE == 1 && acc) {
                const int r = count + __popcll(m & lane_mask_lt());
                if (r < NB_LMAX) s_key[MODE ? wv : 0][r] = key;
            }
            count += __popcll(m);
        }
        wave_sync();
    }
    if (MODE == 0) {
        if (lane == 0) {
            counts[i] = count;
            atomicMax(maxcount, count);
        }
        return;
    }
    // fill: rank sort of the keys, write in reference order (an emptied list means overflow)
    const int base = nb_off[i];
    const int n = min(count < NB_LMAX ? count : NB_LMAX, nb_off[i + 1] - base);
    wave_sync();
    for (int e = lane; e < n; e += 64) {
        const long long k = s_key[MODE ? wv : 0][e];
        int rank = 0;
        for (int f = 0; f < n; ++f) rank += (s_key[MODE ? wv : 0][f] < k);
        const long long jl = k / nimg;
        long long code = k - jl * nimg;
        const int c2 = (int)(code % nz) - S.cf[2];
        code /= nz;
        const int c1 = (int)(code % ny) - S.cf[1];
        const int c0 = (int)(code / ny) - S.cf[0];
        const int j = (int)jl + S.atom_off;
        const int p = base + rank;
        nb_j[p] = j;
        nb_src[p] = i;
        nb_img[p] = img_pack(c0, c1, c2);
        // r_IJ = posn(i) - (posn(j) + R*img)  (nnp.cpp:363), FP64, stored as FP32
        double rv[3];
        for (int d = 0; d < 3; ++d) {
            const double sh = S.R[d] * c0 + S.R[d + 3] * c1 + S.R[d + 6] * c2;
            rv[d] = pos[3 * i + d] - (pos[3 * j + d] + sh);
        }
        const double r = sqrt(rv[0] * rv[0] + rv[1] * rv[1] + rv[2] * rv[2]);
        nb_d[p] = make_float4((float)rv[0], (float)rv[1], (float)rv[2], (float)r);
    }
}

// thread per pair: index of the reverse pair (j -> i, -n) in j's list (binary search on the key)
__global__ void k_nb_reverse(const int* __restrict__ total, const int* __restrict__ nb_src, const int* __restrict__ nb_j,
                             const int* __restrict__ nb_img, const int* __restrict__ nb_off,
                             const int* __restrict__ struc_of, const StrucInfo* __restrict__ sinfo,
                             int* __restrict__ nb_rev, int* __restrict__ err) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= *total) return;
    const int i = nb_src[p], j = nb_j[p];
    const StrucInfo& S = sinfo[struc_of[i]];
    const int ny = 2 * S.cf[1] + 1, nz = 2 * S.cf[2] + 1;
    const long long nimg = (long long)(2 * S.cf[0] + 1) * ny * nz;
    int a, b, c;
    img_unpack(nb_img[p], a, b, c);
    const long long want = (long long)(i - S.atom_off) * nimg +
                           ((long long)(-a + S.cf[0]) * ny + (-b + S.cf[1])) * nz + (-c + S.cf[2]);
    int lo = nb_off[j], hi = nb_off[j + 1] - 1;
    int found = -1;
    while (lo <= hi) {
        const int mid = (lo + hi) >> 1;
        int x, y, z;
        img_unpack(nb_img[mid], x, y, z);
        const long long k = (long long)(nb_j[mid] - S.atom_off) * nimg +
                            ((long long)(x + S.cf[0]) * ny + (y + S.cf[1])) * nz + (z + S.cf[2]);
        if (k == want) { found = mid; break; }
        if (k < want) lo = mid + 1; else hi = mid - 1;
    }
    nb_rev[p] = found < 0 ? p : found;
    if (found < 0) atomicAdd(err, 1);
}

// thread per offset entry: if the new list exceeds the pair capacity or an atom exceeds the LDS
// neighbour slots, every list is emptied and the overflow flag is raised.  The total is read from the
// scan's own copy, which this kernel does not modify.
__global__ void k_nb_check(int natoms, int* __restrict__ nb_off, const int* __restrict__ maxcount,
                           const int* __restrict__ total, int pcap, int slots, int* __restrict__ bad) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i > natoms) return;
    if (*total > pcap || *maxcount > slots) {
        nb_off[i] = 0;
        if (i == 0) *bad = 1;
    }
}

