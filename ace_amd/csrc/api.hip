// api.hip -- the C-ABI (include/ace_hip.h): context, model tables, structure upload, neighbour-list
// build, batched energy/force pass, the plain FP64 port, and per-kernel HIP-event timing.
// The kernel sources are included below; the basis kernels are instantiated in separate translation
// units (basis_inst.hip) and reached through basis_table.hpp.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "ace_common.hpp"
#include "neighbors.hip"
#include "basis.hip"
#include "basis_table.hpp"
#include "mlp.hip"
#include "reduce.hip"
#include "plain.hip"

#define NB_CELL_DIV 2   // cells per list radius along each perpendicular height (binned search)
#include "train.hip"
#include "md.hip"

namespace {

struct HSpecies {
    int K = 0, nR = 0, L = 0, act = 0;
    double c = 1, outw = 1, outb = 0, E0 = 0;
    std::vector<int> r_form, r_cut, r_n;
    std::vector<double> r_rc;
    std::vector<std::vector<double>> r_rs, r_eta;
    std::vector<int> a_form, a_cut, a_n;
    std::vector<double> a_rc;
    std::vector<std::vector<double>> a_eta, a_zeta;
    std::vector<std::vector<int>> a_lambda;
    std::vector<int> width;
    std::vector<double> inpw, inpb, params;
    int nparams() const {
        int s = 0;
        for (int l = 0; l < L; ++l) s += width[l + 1] + width[l + 1] * width[l];
        return s;
    }
};

struct ProfRec {
    std::string name;
    hipEvent_t a, b;
};

}  // namespace

struct ace_ctx {
    std::string err;
    hipStream_t stream = nullptr;
    // model
    int nt = 0;
    double rc = 0;
    std::vector<HSpecies> sp;
    int kmax = 0, ldg = 0, maxA = 0, maxR = 0, ngmax = 1, ngk = 1, nzmax = 1;
    int nrp = 1;                     // most radial functions of one species
    bool all_canon = true;
    int lsplit = 0;                  // level split of the two eta groups' canonical levels (0: none)
    bool lsplit_on = true;           // ace_set_level_split
    bool fgather = true;             // ace_set_force_gather: energy/force passes gather per-pair gradients
    DevModel hdm;
    DBuf<DevModel> d_model;
    DBuf<float> d_r_rs, d_r_eta, d_a_lamf, d_a_zeta, d_inpw, d_inpb, d_wpad;
    DBuf<uint16_t> d_w0bf;           // BF16 split copies of every species' first-layer weights
    // pinned staging of the layouts above (reused; the copy out of it is fenced by stage_done)
    char* h_stage = nullptr;
    size_t h_stage_bytes = 0;
    hipEvent_t stage_done = nullptr;
    int wp_total = 0, bf_total = 0;
    std::vector<int> pmap;           // device-layout indices of every parameter (param_map)
    DBuf<int> d_a_zint, d_a_grp, d_a_sel, d_a_cidx;
    bool plain_ok = false;
    std::string plain_why;
    PlainModel hpm;
    DBuf<PlainModel> d_pmodel;
    DBuf<double> d_r_rs64, d_r_eta64, d_a_eta64, d_a_zeta64, d_a_lamf64, d_nn64, d_inpw64, d_inpb64;
    // ace_autotune's choices (-1: the size rule), valid for the batch size and row capacity measured on
    int tune_fused[3] = {-1, -1, -1};   // per precision: fused kernel with 4 (1) or 2 (2) waves per atom, or
                                        // symm -> network -> adjoint (0)
    int tune_stage[3] = {-1, -1, -1};   // per precision: network weights staged in LDS (1) or read from memory
    int tune_mlpw[3] = {0, 0, 0};       // per precision: waves per network tile (0: the default, 4)
    int pin_chunk[3] = {0, 0, 0};       // ace_set_tile_shape: fixed chunk weight w / waves per network tile
    int pin_mlpw[3] = {0, 0, 0};        // (0: tuned or default)
    double tune_mlpw_ms[3][3] = {{-1, -1, -1}, {-1, -1, -1}, {-1, -1, -1}};   // timings for 1, 2, 4 waves
    int tune_waves[3] = {-1, -1, -1};   // per precision: waves per atom of the split path's kernels (4 or 2)
    int tune_chunk[3] = {-1, -1, -1};   // per precision: the angular chunk weight w (1, 3 or 8; -1: default 3)
    double tune_chunk_ms[3][3] = {{-1, -1, -1}, {-1, -1, -1}, {-1, -1, -1}};   // its timings, w = 1, 3, 8
    long long tune_N = -1;
    int tune_slots = -1;
    bool pinned[3] = {false, false, false};   // ace_set_path: the choice holds for every batch
    bool params_dirty = true;        // network parameters changed: matrix-core layouts to upload
    bool scale_dirty = true;         // input scaling changed
    bool nn64_dirty = true;          // FP64 parameter copy of the plain port is stale
    int precision = ACE_PREC_FP32;   // default of ace_symm / ace_train_step
    bool basis_f64 = false;          // precision of the basis kernels of the current call
    unsigned long long seq = 0;      // advances with every device pass and every change of positions,
                                     // parameters, input scaling or precision (ace_backward's validity check)
    // structures
    int S = 0, N = 0;
    std::vector<int> natoms, atom_off, type_h, row_of_h, struc_of_h;
    std::vector<double> cell;
    DBuf<int> d_type, d_struc_of, d_atom_off, d_natoms, d_row_of, d_atom_of_row;
    DBuf<double> d_pos, d_cell;
    std::vector<int4> tiles_h;
    DBuf<int4> d_tiles;
    int tm = 32;
    DBuf<float> d_G, d_dEdG;
    DBuf<double> d_e_atom, d_F, d_E, d_vir;
    // fixed-point accumulators, two buffers: a gradient launch adds into buffer fbuf and zeroes the other
    DBuf<unsigned long long> d_Ffix[2], d_Efix[2], d_Wfix[2];
    int fbuf = 0;
    // neighbours
    bool nb_valid = false;
    double nb_rc = 0;
    std::vector<StrucInfo> sinfo_h, sinfo_up;   // host tables; the copy last uploaded
    DBuf<StrucInfo> d_sinfo;
    int ncells = 0;
    DBuf<int> d_cell_of, d_wrap_of, d_slot_of, d_cell_start, d_nb_count, d_nb_off, d_scan_blk, d_scal;
    DBuf<double4> d_cpos;      // cell-ordered positions
    DBuf<int2> d_cinfo;        // cell-ordered (atom, wrap)
    DBuf<long long> d_nb_keys; // per-atom rows of sorted neighbour keys
    DBuf<double> d_gin;        // gathered energy/force pass: per-entry gradients, rows like the keys
    DBuf<double> d_fown;       // and each atom's own share
    DBuf<int> d_rev;           // reverse entries of the key-row entries (written by the symmetry pass)
    long long nb_builds = 0;   // neighbour builds so far
    long long rev_build = -1;  // the build whose reverse entries d_rev holds
    long long npairs = 0;       // pairs of the current list (set when the CSR arrays are built)
    bool csr_valid = false;     // CSR arrays (export, unbatched port) match the key rows
    int slots_cap = 32;         // neighbour slots per atom held in LDS by the basis kernels
    bool nb_async_ok = false;   // capacities known: rebuilds need no host round trip
    bool nb_same_layout = false;
    bool nb_pending = false;    // a rebuild's overflow / symmetry flags are not checked yet
    bool nb_brute = false;      // the last build tested every atom of each structure (k_nb_brute)
    int nb_over_idx = 5;        // d_scal index of the last build's list-overflow flag
    int nb_parity = 0;          // which of d_scal[6], d_scal[7] the next brute-force build raises
    double nb_skin = 0;
    DBuf<int> d_nb_j, d_nb_img;
    // plain port scratch
    DBuf<double> d_G64;
    // training
    TrainState tr;
    // molecular dynamics: velocities, per-species 1/(m mvv2e) then m mvv2e, per-step energies, step counter,
    // and the captured two-step graphs (by the accumulator buffer the first step adds into)
    DBuf<double> d_vel, d_minv;
    DBuf<unsigned long long> d_md_rec;
    DBuf<unsigned int> d_md_ctr;
    bool md_ready = false;
    hipGraphExec_t md_exec[2] = {nullptr, nullptr};
    long long md_sig[2][8] = {};
    unsigned long long gen = 0;   // counts device reallocations: captured graphs hold raw pointers
    // device of the context and its address of the no-allocation scratch array (ace_create)
    int device = -1;
    unsigned char* scratch = nullptr;
    // profiling
    bool prof = false;
    std::vector<ProfRec> pending;
    std::map<std::string, std::pair<double, long long>> acc;
};

// Scratch memory of ace_energy_forces_noalloc: fixed-point accumulators, per-atom energies, network rows and
// the virial when the caller passes none.  One copy per GPU, held until the process ends.  In host code
// `scratch` names a host-side stand-in, not device memory: its device address is looked up once per device
// (scratch_on, from ace_create, with that device current) and only that address is used.
#define ACE_NOALLOC_BYTES (64ull << 20)
static __device__ __align__(256) unsigned char scratch[ACE_NOALLOC_BYTES];

namespace {

std::mutex g_scratch_mu;
unsigned char* g_scratch_ptr[64] = {};

// the current device's address of `scratch` (nullptr if it cannot be looked up)
unsigned char* scratch_on(int dev) {
    if (dev < 0 || dev >= 64) return nullptr;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess) return nullptr;
        g_scratch_ptr[dev] = (unsigned char*)p;
    }
    return g_scratch_ptr[dev];
}

// inverse of a 3x3 matrix, both column-major (Rinv[r + 3 c] = (R^-1)(r, c)); R must be non-singular
void mat3_inverse(const double* R, double* Rinv) {
    const double m[3][3] = {{R[0], R[3], R[6]}, {R[1], R[4], R[7]}, {R[2], R[5], R[8]}};
    const double det = m[0][0] * (m[1][1] * m[2][2] - m[1][2] * m[2][1]) - m[0][1] * (m[1][0] * m[2][2] - m[1][2] * m[2][0]) +
                       m[0][2] * (m[1][0] * m[2][1] - m[1][1] * m[2][0]);
    const double id = 1.0 / det;
    double inv[3][3];
    inv[0][0] = (m[1][1] * m[2][2] - m[1][2] * m[2][1]) * id;
    inv[0][1] = (m[0][2] * m[2][1] - m[0][1] * m[2][2]) * id;
    inv[0][2] = (m[0][1] * m[1][2] - m[0][2] * m[1][1]) * id;
    inv[1][0] = (m[1][2] * m[2][0] - m[1][0] * m[2][2]) * id;
    inv[1][1] = (m[0][0] * m[2][2] - m[0][2] * m[2][0]) * id;
    inv[1][2] = (m[0][2] * m[1][0] - m[0][0] * m[1][2]) * id;
    inv[2][0] = (m[1][0] * m[2][1] - m[1][1] * m[2][0]) * id;
    inv[2][1] = (m[0][1] * m[2][0] - m[0][0] * m[2][1]) * id;
    inv[2][2] = (m[0][0] * m[1][1] - m[0][1] * m[1][0]) * id;
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) Rinv[r + 3 * c] = inv[r][c];
}

int fail(ace_ctx* c, const std::string& msg) {
    if (c) c->err = msg;
    return -1;
}
int fail_hip(ace_ctx* c, const char* what, hipError_t e) {
    return fail(c, std::string(what) + ": " + hipGetErrorString(e));
}
#define HC(x)                                                         \
    do {                                                              \
        hipError_t e_ = (x);                                          \
        if (e_ != hipSuccess) return fail_hip(ctx, #x, e_);          \
    } while (0)

template <class T>
int ensure(ace_ctx* ctx, DBuf<T>& b, size_t n) {
    if (n == 0) n = 1;
    if (b.cap >= n) return 0;
    if (b.p) HC(hipFree(b.p));
    b.p = nullptr;
    b.cap = 0;
    ++ctx->gen;
    HC(hipMalloc((void**)&b.p, n * sizeof(T)));
    b.cap = n;
    return 0;
}
template <class T>
void release(DBuf<T>& b) {
    if (b.p) (void)hipFree(b.p);
    b.p = nullptr;
    b.cap = 0;
}
template <class T>
int upload(ace_ctx* ctx, DBuf<T>& b, const T* h, size_t n) {
    if (ensure(ctx, b, n)) return -1;
    if (n) HC(hipMemcpyAsync(b.p, h, n * sizeof(T), hipMemcpyHostToDevice, ctx->stream));
    return 0;
}

// kernel launch with optional HIP-event timing on the context's stream
template <class F>
int timed(ace_ctx* ctx, const char* name, F&& launch) {
    ProfRec r;
    if (ctx->prof) {
        r.name = name;
        HC(hipEventCreate(&r.a));
        HC(hipEventCreate(&r.b));
        HC(hipEventRecord(r.a, ctx->stream));
    }
    launch();
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail_hip(ctx, name, e);
    if (ctx->prof) {
        HC(hipEventRecord(r.b, ctx->stream));
        ctx->pending.push_back(r);
    }
    return 0;
}

int flush_prof(ace_ctx* ctx) {
    if (ctx->pending.empty()) return 0;
    HC(hipStreamSynchronize(ctx->stream));
    for (auto& r : ctx->pending) {
        float ms = 0;
        HC(hipEventElapsedTime(&ms, r.a, r.b));
        auto& a = ctx->acc[r.name];
        a.first += ms;
        a.second += 1;
        (void)hipEventDestroy(r.a);
        (void)hipEventDestroy(r.b);
    }
    ctx->pending.clear();
    return 0;
}

inline int cdiv(long long a, long long b) { return (int)((a + b - 1) / b); }

int exclusive_scan(ace_ctx* ctx, const int* in, int* out, int n) {
    const int nb = cdiv(n > 0 ? n : 1, SCAN_BLOCK);
    if (ensure(ctx, ctx->d_scan_blk, nb + 1)) return -1;
    int* blk = ctx->d_scan_blk.p;
    int* total = ctx->d_scal.p + 1;
    if (n <= SCAN_BLOCK * SCAN_RUN)
        return timed(ctx, "scan", [&] { k_scan_single<<<1, SCAN_BLOCK, 0, ctx->stream>>>(n, in, out, total); });
    if (timed(ctx, "scan_local", [&] { k_scan_local<<<nb, SCAN_BLOCK, 0, ctx->stream>>>(n, in, out, blk); })) return -1;
    if (timed(ctx, "scan_sums", [&] { k_scan_sums<<<1, SCAN_BLOCK, 0, ctx->stream>>>(nb, blk, total, n, out); })) return -1;
    if (timed(ctx, "scan_add", [&] { k_scan_add<<<nb, SCAN_BLOCK, 0, ctx->stream>>>(n, out, blk); })) return -1;
    return 0;
}

// ---- model ------------------------------------------------------------------------------------

int parse_model(ace_ctx* ctx, const ace_model_desc* m) {
    if (!m) return fail(ctx, "ace_create: null model");
    if (m->ntypes < 1 || m->ntypes > ACE_MAX_TYPES) return fail(ctx, "ace_create: ntypes out of range [1,8]");
    if (!(m->rc > 0)) return fail(ctx, "ace_create: invalid cutoff radius");
    const int nt = m->ntypes, nab = nt * (nt + 1) / 2;
    ctx->nt = nt;
    ctx->rc = m->rc;
    ctx->sp.assign(nt, HSpecies());
    for (int I = 0; I < nt; ++I) {
        HSpecies& s = ctx->sp[I];
        for (int J = 0; J < nt; ++J) {
            const int b = I * nt + J;
            s.r_form.push_back(m->r_form[b]);
            s.r_cut.push_back(m->r_cut[b]);
            s.r_rc.push_back(m->r_rc[b]);
            s.r_n.push_back(m->r_n[b]);
            const double* rs = m->r_rs + m->r_off[b];
            const double* et = m->r_eta + m->r_off[b];
            s.r_rs.emplace_back(rs, rs + m->r_n[b]);
            s.r_eta.emplace_back(et, et + m->r_n[b]);
            if (m->r_n[b] < 0) return fail(ctx, "ace_create: negative radial block size");
            if (m->r_n[b] > 0 && (m->r_form[b] < ACE_RAD_GAUSSIAN || m->r_form[b] > ACE_RAD_LOGCOSH2))
                return fail(ctx, "ace_create: invalid radial function name");
            if (m->r_n[b] > 0 && (m->r_cut[b] < ACE_CUT_STEP || m->r_cut[b] > ACE_CUT_POLY3))
                return fail(ctx, "ace_create: invalid cutoff name");
            if (m->r_n[b] > 0 && !(m->r_rc[b] > 0)) return fail(ctx, "ace_create: invalid basis cutoff");
        }
        for (int b = 0; b < nab; ++b) {
            const int q = I * nab + b;
            s.a_form.push_back(m->a_form[q]);
            s.a_cut.push_back(m->a_cut[q]);
            s.a_rc.push_back(m->a_rc[q]);
            s.a_n.push_back(m->a_n[q]);
            const int o = m->a_off[q], n = m->a_n[q];
            s.a_eta.emplace_back(m->a_eta + o, m->a_eta + o + n);
            s.a_zeta.emplace_back(m->a_zeta + o, m->a_zeta + o + n);
            s.a_lambda.emplace_back(m->a_lambda + o, m->a_lambda + o + n);
            if (n < 0) return fail(ctx, "ace_create: negative angular block size");
            if (n > 0 && (m->a_form[q] < ACE_ANG_GAUSS || m->a_form[q] > ACE_ANG_STUDENT5))
                return fail(ctx, "ace_create: invalid angular function name");
            if (n > 0 && (m->a_cut[q] < ACE_CUT_STEP || m->a_cut[q] > ACE_CUT_POLY3))
                return fail(ctx, "ace_create: invalid cutoff name");
            if (n > 0 && !(m->a_rc[q] > 0)) return fail(ctx, "ace_create: invalid basis cutoff");
        }
        s.L = m->nn_nlayer[I];
        if (s.L < 1 || s.L > ACE_MAX_LAYERS) return fail(ctx, "ace_create: nlayer out of range");
        const int* w = m->nn_width + I * (ACE_MAX_LAYERS + 1);
        s.width.assign(w, w + s.L + 1);
        for (int l = 0; l <= s.L; ++l)
            if (s.width[l] < 1) return fail(ctx, "ace_create: invalid layer width");
        if (s.width[s.L] != 1) return fail(ctx, "ace_create: the output layer must have width 1");
        s.act = m->nn_act[I];
        if (s.act != ACE_ACT_LINEAR && s.act != ACE_ACT_TANH && s.act != ACE_ACT_SIGMOID && s.act != ACE_ACT_SWISH &&
            s.act != ACE_ACT_SOFTPLUS)
            return fail(ctx, "ace_create: activation not supported");
        s.c = m->nn_c[I];
        s.outw = m->outw[I];
        s.outb = m->outb[I];
        s.E0 = m->E0[I];
        int K = 0;
        for (int J = 0; J < nt; ++J) K += s.r_n[J];
        s.nR = K;
        for (int b = 0; b < nab; ++b) K += s.a_n[b];
        s.K = K;
        if (s.width[0] != K) return fail(ctx, "ace_create: network input width differs from the basis size");
        s.inpw.assign(m->inpw + m->in_off[I], m->inpw + m->in_off[I] + K);
        s.inpb.assign(m->inpb + m->in_off[I], m->inpb + m->in_off[I] + K);
        s.params.assign(m->params + m->par_off[I], m->params + m->par_off[I] + s.nparams());
    }
    return 0;
}

// matrix-core copy of the network parameters (layout in DevSpecies::wl_off / bl_off)
// FP32 -> BF16 bits, round to nearest even (finite inputs)
__host__ __device__ inline uint16_t bf16_rne(float x) {
    uint32_t u = __builtin_bit_cast(uint32_t, x);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__host__ __device__ inline float bf16_value(uint16_t h) { return __builtin_bit_cast(float, (uint32_t)h << 16); }

// Where each parameter lives in the device layouts, 5 ints per parameter (species concatenated in type
// order, each species' parameters in the ANN order [biases][weights] of its layers): the FP32 matrix-core
// index in wpad, then for first-layer weights the BF16 high / low indices of the forward [o][k] and the
// backward [k][o] copies in w0bf (-1 elsewhere).
const std::vector<int>& param_map(ace_ctx* ctx) {
    if (!ctx->pmap.empty()) return ctx->pmap;
    std::vector<int>& map = ctx->pmap;
    int ntot = 0;
    for (int I = 0; I < ctx->nt; ++I) ntot += ctx->sp[I].nparams();
    map.assign(5 * (size_t)ntot, -1);
    int q0 = 0;
    for (int I = 0; I < ctx->nt; ++I) {
        const HSpecies& s = ctx->sp[I];
        const DevSpecies& d = ctx->hdm.sp[I];
        const int K = s.width[0], n1 = s.width[1];
        const int fwd = ((n1 + 15) & ~15) * ((K + 31) & ~31), bwd = ((K + 15) & ~15) * ((n1 + 31) & ~31);
        for (int l = 0; l < s.L; ++l) {
            const int ni = s.width[l], no = s.width[l + 1];
            for (int k = 0; k < ni; ++k)
                for (int o = 0; o < no; ++o) {
                    int* e = &map[5 * (size_t)(q0 + d.woff[l] + o + k * no)];
                    e[0] = d.wp_off + d.wl_off[l] + k * mlp_ld(no) + o;
                    if (l == 0) {
                        const int ef = d.bf_off + o * ((K + 31) & ~31) + k;
                        const int eb = d.bf_off + 2 * fwd + k * ((n1 + 31) & ~31) + o;
                        e[1] = ef; e[2] = ef + fwd; e[3] = eb; e[4] = eb + bwd;
                    }
                }
            for (int o = 0; o < no; ++o) map[5 * (size_t)(q0 + d.boff[l] + o)] = d.wp_off + d.bl_off[l] + o;
        }
        q0 += s.nparams();
    }
    return map;
}

// The matrix-core layouts of every network (DevSpecies::wp_off / wl_off / bl_off, and the BF16 split copies
// of W_0 at bf_off) built in pinned host memory and copied asynchronously; the buffer is reused once the
// previous copy out of it has completed.  Parameters change every training step, so this is on that path.
int upload_wpad(ace_ctx* ctx) {
    const size_t nw = std::max(ctx->wp_total, 4), nb = std::max(ctx->bf_total, 8);
    const size_t bytes = nw * sizeof(float) + nb * sizeof(uint16_t);
    if (ctx->stage_done) HC(hipEventSynchronize(ctx->stage_done));
    else HC(hipEventCreateWithFlags(&ctx->stage_done, hipEventDisableTiming));
    if (ctx->h_stage_bytes < bytes) {
        if (ctx->h_stage) HC(hipHostFree(ctx->h_stage));
        ctx->h_stage = nullptr;
        ctx->h_stage_bytes = 0;
        HC(hipHostMalloc((void**)&ctx->h_stage, bytes, hipHostMallocDefault));
        ctx->h_stage_bytes = bytes;
    }
    if (ensure(ctx, ctx->d_wpad, nw) || ensure(ctx, ctx->d_w0bf, nb)) return -1;
    float* w = (float*)ctx->h_stage;
    uint16_t* b = (uint16_t*)(ctx->h_stage + nw * sizeof(float));
    memset(ctx->h_stage, 0, bytes);
    const std::vector<int>& map = param_map(ctx);
    size_t q = 0;
    for (int I = 0; I < ctx->nt; ++I)
        for (const double v : ctx->sp[I].params) {
            const int* e = &map[5 * q++];
            const float x = (float)v;
            w[e[0]] = x;
            if (e[1] >= 0) {
                // BF16 high/low parts of W_0 in the two layouts the BF16 first layer reads
                const uint16_t hi = bf16_rne(x), lo = bf16_rne(x - bf16_value(hi));
                b[e[1]] = hi; b[e[2]] = lo; b[e[3]] = hi; b[e[4]] = lo;
            }
        }
    HC(hipMemcpyAsync(ctx->d_wpad.p, w, nw * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
    HC(hipMemcpyAsync(ctx->d_w0bf.p, b, nb * sizeof(uint16_t), hipMemcpyHostToDevice, ctx->stream));
    HC(hipEventRecord(ctx->stage_done, ctx->stream));
    return 0;
}

// BasisA::init run grouping (basis_angular.cpp:178-205), merged over the blocks of one species
int build_tables(ace_ctx* ctx) {
    const int nt = ctx->nt, nab = nt * (nt + 1) / 2;
    DevModel& dm = ctx->hdm;
    memset(&dm, 0, sizeof(dm));
    dm.ntypes = nt;
    std::vector<float> r_rs, r_eta, a_lamf, a_zeta, prm, iw, ib;
    std::vector<int> a_zint, a_grp, a_sel;
    int kmax = 0, maxA = 1, maxR = 1, ngmax = 1, wp_total = 0, bf_total = 0;
    for (int I = 0; I < nt; ++I) {
        const HSpecies& s = ctx->sp[I];
        DevSpecies& d = dm.sp[I];
        d.K = s.K;
        d.nR = s.nR;
        kmax = std::max(kmax, s.K);
        int goff = 0;
        double rcm = 0;
        for (int J = 0; J < nt; ++J) {
            d.r_form[J] = s.r_form[J];
            d.r_cut[J] = s.r_cut[J];
            d.r_n[J] = s.r_n[J];
            d.r_rc[J] = (float)s.r_rc[J];
            d.r_rc64[J] = s.r_rc[J];
            d.r_goff[J] = goff;
            d.r_poff[J] = (int)r_rs.size();
            goff += s.r_n[J];
            for (int q = 0; q < s.r_n[J]; ++q) {
                r_rs.push_back((float)s.r_rs[J][q]);
                r_eta.push_back((float)s.r_eta[J][q]);
            }
            if (s.r_n[J] > 0) rcm = std::max(rcm, s.r_rc[J]);
            maxR = std::max(maxR, s.r_n[J]);
        }
        // angular: one cutoff per species, distinct etas merged into groups
        int acut = -1;
        double arc = 0;
        std::vector<double> groups;
        for (int b = 0; b < nab; ++b) {
            if (s.a_n[b] == 0) continue;
            if (acut < 0) { acut = s.a_cut[b]; arc = s.a_rc[b]; }
            else if (acut != s.a_cut[b] || arc != s.a_rc[b])
                return fail(ctx, "ace_create: the GPU path needs one angular cutoff per species");
        }
        d.a_cut = acut < 0 ? ACE_CUT_COS : acut;
        d.a_rc = (float)(acut < 0 ? ctx->rc : arc);
        d.a_rc64 = acut < 0 ? ctx->rc : arc;
        if (acut >= 0) rcm = std::max(rcm, arc);
        d.rc_max = (float)std::min(rcm, ctx->rc);
        for (int b = 0; b < nab; ++b) {
            d.a_form[b] = s.a_form[b];
            d.a_n[b] = s.a_n[b];
            d.a_goff[b] = goff;
            d.a_poff[b] = (int)a_lamf.size();
            goff += s.a_n[b];
            maxA = std::max(maxA, s.a_n[b]);
            // run grouping within the block: consecutive etas closer than 1e-6 share the first one
            double run_eta = 0;
            for (int q = 0; q < s.a_n[b]; ++q) {
                const double e = s.a_eta[b][q];
                if (q == 0 || std::fabs(e - s.a_eta[b][q - 1]) >= 1.0e-6) run_eta = e;
                int g = -1;
                for (size_t t = 0; t < groups.size(); ++t)
                    if (groups[t] == run_eta) g = (int)t;
                if (g < 0) {
                    if ((int)groups.size() >= ACE_MAXG)
                        return fail(ctx, "ace_create: more than 4 distinct angular eta values in one species");
                    groups.push_back(run_eta);
                    g = (int)groups.size() - 1;
                }
                a_lamf.push_back((float)(0.5 * s.a_lambda[b][q]));
                a_zeta.push_back((float)s.a_zeta[b][q]);
                const double z = s.a_zeta[b][q];
                a_zint.push_back((z == std::floor(z) && z >= 1 && z <= 64) ? (int)z : -1);
                a_grp.push_back(g);
                if (!(z > 0)) return fail(ctx, "ace_create: zeta must be positive");
                // power-of-two zeta <= 64 with lambda = +-1: one pick from the per-triplet power cache
                int sel = -1;
                const int zi = (z == std::floor(z) && z >= 1 && z <= 64) ? (int)z : 0;
                if (zi && (zi & (zi - 1)) == 0 && std::abs(s.a_lambda[b][q]) == 1) {
                    int k = 0;
                    while ((1 << k) < zi) ++k;
                    sel = k + (s.a_lambda[b][q] < 0 ? 7 : 0);
                }
                a_sel.push_back(sel);
            }
        }
        d.ng = (int)groups.size();
        for (int g = 0; g < d.ng; ++g) {
            d.ietar2[g] = (float)(1.0 / (groups[g] * groups[g]));
            d.ietar2_64[g] = 1.0 / (groups[g] * groups[g]);
        }
        ngmax = std::max(ngmax, d.ng);
        // network
        d.L = s.L;
        for (int l = 0; l <= s.L; ++l) d.width[l] = s.width[l];
        d.act = s.act;
        d.c = (float)s.c;
        int cnt = 0;
        for (int l = 0; l < s.L; ++l) { d.boff[l] = cnt; cnt += s.width[l + 1]; }
        for (int l = 0; l < s.L; ++l) { d.woff[l] = cnt; cnt += s.width[l + 1] * s.width[l]; }
        d.ldp = cnt;
        d.poff = (int)prm.size();
        {
            int off = 0;
            // W_l as [k][o]: rows padded to a multiple of 16, row stride = mlp_ld(w_{l+1}); zero padding
            for (int l = 0; l < s.L; ++l) {
                d.wl_off[l] = off;
                off += ((s.width[l] + 15) & ~15) * mlp_ld(s.width[l + 1]);
            }
            for (int l = 0; l < s.L; ++l) { d.bl_off[l] = off; off += (s.width[l + 1] + 15) & ~15; }
            d.wp_n = (off + 3) & ~3;
            d.wp_off = wp_total;
            wp_total += d.wp_n;
            const int K = s.width[0], n1 = s.width[1];
            d.bf_off = bf_total;
            bf_total += 2 * ((n1 + 15) & ~15) * ((K + 31) & ~31) + 2 * ((K + 15) & ~15) * ((n1 + 31) & ~31);
        }
        for (double v : s.params) prm.push_back((float)v);
        d.in_off = (int)iw.size();
        for (int k = 0; k < s.K; ++k) {
            iw.push_back((float)s.inpw[k]);
            ib.push_back((float)s.inpb[k]);
        }
        d.outw = (float)s.outw;
        d.outb = (float)s.outb;
        d.E0 = s.E0;
    }
    // canonical angular blocks: every function lambda = +-1, zeta = 2^k (k <= 6) and one function per
    // (sign, level, eta group); their sums become a register outer product in the basis kernel
    const int ngk = ngmax <= 1 ? 1 : (ngmax <= 2 ? 2 : 4);
    std::vector<int> a_cidx(a_sel.size(), 0);
    bool all_canon = true;
    int nzmax_all = 1;
    for (int I = 0; I < nt; ++I) {
        DevSpecies& d = dm.sp[I];
        int nz = 1;
        for (int b = 0; b < nab; ++b) {
            bool canon = d.a_n[b] > 0 && d.a_form[b] == ACE_ANG_GAUSS;   // the others are not separable
            std::vector<char> used(14 * ngk, 0);
            for (int q = 0; q < d.a_n[b] && canon; ++q) {
                const int fi = d.a_poff[b] + q, sel = a_sel[fi];
                if (sel < 0) { canon = false; break; }
                const int c = sel * ngk + a_grp[fi];
                if (used[c]) { canon = false; break; }
                used[c] = 1;
                a_cidx[fi] = c;
            }
            d.a_canon[b] = canon ? 1 : 0;
            if (d.a_n[b] > 0 && !canon) all_canon = false;
            if (canon)
                for (int q = 0; q < d.a_n[b]; ++q) nz = std::max(nz, a_sel[d.a_poff[b] + q] % 7 + 1);
        }
        d.nz = nz;
        nzmax_all = std::max(nzmax_all, nz);
    }
    ctx->nzmax = nzmax_all;
    ctx->ngk = ngk;
    ctx->all_canon = all_canon;
    // two eta groups whose canonical levels do not overlap -- group 0 below level s, group 1 from s on --
    // take the kernels that form only those (level, group) products
    ctx->lsplit = 0;
    if (ngk == 2) {
        int top0 = -1, low1 = 99;
        for (int I = 0; I < nt; ++I) {
            const DevSpecies& d = dm.sp[I];
            for (int b = 0; b < nab; ++b) {
                if (!d.a_canon[b]) continue;
                for (int q = 0; q < d.a_n[b]; ++q) {
                    const int fi = d.a_poff[b] + q, k = a_sel[fi] % 7;
                    if (a_grp[fi] == 0) top0 = std::max(top0, k);
                    else low1 = std::min(low1, k);
                }
            }
        }
        if (top0 >= 0 && low1 < 99 && top0 < low1) ctx->lsplit = top0 + 1;
    }
    dm.kmax = kmax;
    dm.ng_max = ngmax;
    ctx->kmax = kmax;
    ctx->ldg = (kmax + 3) & ~3;
    ctx->maxA = maxA;
    ctx->maxR = maxR;
    ctx->nrp = 1;
    for (int I = 0; I < nt; ++I) ctx->nrp = std::max(ctx->nrp, ctx->sp[I].nR);
    ctx->ngmax = ngmax;
    if (upload(ctx, ctx->d_model, &dm, 1)) return -1;
    if (upload(ctx, ctx->d_r_rs, r_rs.data(), r_rs.size())) return -1;
    if (upload(ctx, ctx->d_r_eta, r_eta.data(), r_eta.size())) return -1;
    if (upload(ctx, ctx->d_a_lamf, a_lamf.data(), a_lamf.size())) return -1;
    if (upload(ctx, ctx->d_a_zeta, a_zeta.data(), a_zeta.size())) return -1;
    if (upload(ctx, ctx->d_a_zint, a_zint.data(), a_zint.size())) return -1;
    if (upload(ctx, ctx->d_a_grp, a_grp.data(), a_grp.size())) return -1;
    if (upload(ctx, ctx->d_a_sel, a_sel.data(), a_sel.size())) return -1;
    if (upload(ctx, ctx->d_a_cidx, a_cidx.data(), a_cidx.size())) return -1;
    if (upload(ctx, ctx->d_inpw, iw.data(), iw.size())) return -1;
    if (upload(ctx, ctx->d_inpb, ib.data(), ib.size())) return -1;
    ctx->wp_total = wp_total;
    ctx->bf_total = bf_total;
    if (upload_wpad(ctx)) return -1;

    // FP64 tables of the plain port
    ctx->plain_ok = true;
    PlainModel& pm = ctx->hpm;
    memset(&pm, 0, sizeof(pm));
    pm.ntypes = nt;
    std::vector<double> rrs, reta, aeta, azeta, alamf, nn, iw64, ib64;
    for (int I = 0; I < nt; ++I) {
        const HSpecies& s = ctx->sp[I];
        PlainSpecies& p = pm.sp[I];
        p.K = s.K; p.nR = s.nR; p.L = s.L; p.act = s.act;
        p.c = s.c; p.outw = s.outw; p.outb = s.outb; p.E0 = s.E0;
        if (s.K > PL_WMAX || s.L > PL_LMAX) { ctx->plain_ok = false; ctx->plain_why = "network too large"; }
        if (s.act != ACE_ACT_TANH && s.act != ACE_ACT_LINEAR) { ctx->plain_ok = false; ctx->plain_why = "activation"; }
        int goff = 0;
        for (int J = 0; J < nt; ++J) {
            p.r_n[J] = s.r_n[J]; p.r_goff[J] = goff; p.r_poff[J] = (int)rrs.size(); p.r_rc[J] = s.r_rc[J];
            goff += s.r_n[J];
            if (s.r_n[J] && (s.r_form[J] != ACE_RAD_GAUSSIAN || s.r_cut[J] != ACE_CUT_COS)) {
                ctx->plain_ok = false; ctx->plain_why = "radial form";
            }
            for (int q = 0; q < s.r_n[J]; ++q) { rrs.push_back(s.r_rs[J][q]); reta.push_back(s.r_eta[J][q]); }
        }
        for (int b = 0; b < nab; ++b) {
            p.a_n[b] = s.a_n[b]; p.a_goff[b] = goff; p.a_poff[b] = (int)aeta.size(); p.a_rc[b] = s.a_rc[b];
            goff += s.a_n[b];
            if (s.a_n[b] && s.a_cut[b] != ACE_CUT_COS) { ctx->plain_ok = false; ctx->plain_why = "angular cutoff"; }
            if (s.a_n[b] && s.a_form[b] != ACE_ANG_GAUSS) { ctx->plain_ok = false; ctx->plain_why = "angular form"; }
            for (int q = 0; q < s.a_n[b]; ++q) {
                aeta.push_back(s.a_eta[b][q]); azeta.push_back(s.a_zeta[b][q]); alamf.push_back(0.5 * s.a_lambda[b][q]);
            }
        }
        for (int l = 0; l <= s.L && l <= ACE_MAXL; ++l) p.width[l] = s.width[l];
        int cnt = 0;
        for (int l = 0; l < s.L; ++l) { p.boff[l] = cnt; cnt += s.width[l + 1]; }
        for (int l = 0; l < s.L; ++l) { p.woff[l] = cnt; cnt += s.width[l + 1] * s.width[l]; }
        p.poff = (int)nn.size();
        nn.insert(nn.end(), s.params.begin(), s.params.end());
        p.in_off = (int)iw64.size();
        iw64.insert(iw64.end(), s.inpw.begin(), s.inpw.end());
        ib64.insert(ib64.end(), s.inpb.begin(), s.inpb.end());
    }
    if (upload(ctx, ctx->d_pmodel, &pm, 1)) return -1;
    if (upload(ctx, ctx->d_r_rs64, rrs.data(), rrs.size())) return -1;
    if (upload(ctx, ctx->d_r_eta64, reta.data(), reta.size())) return -1;
    if (upload(ctx, ctx->d_a_eta64, aeta.data(), aeta.size())) return -1;
    if (upload(ctx, ctx->d_a_zeta64, azeta.data(), azeta.size())) return -1;
    if (upload(ctx, ctx->d_a_lamf64, alamf.data(), alamf.size())) return -1;
    if (upload(ctx, ctx->d_nn64, nn.data(), nn.size())) return -1;
    if (upload(ctx, ctx->d_inpw64, iw64.data(), iw64.size())) return -1;
    if (upload(ctx, ctx->d_inpb64, ib64.data(), ib64.size())) return -1;
    HC(hipStreamSynchronize(ctx->stream));   // host vectors above go out of scope
    ctx->params_dirty = false;
    return 0;
}

// device copies of what changed since the last launch: the matrix-core layouts of the networks, the input
// scaling (FP32 and FP64); the plain port's FP64 parameters are refreshed by the port itself
// parameters updated by the device optimiser, copied back into the host copies (one round trip)
int sync_host_params(ace_ctx* ctx) {
    TrainState& T = ctx->tr;
    if (!T.opt_on || !T.opt_host_stale) return 0;
    std::vector<double> p(T.opt_n);
    HC(hipMemcpyAsync(p.data(), T.d_p.p, T.opt_n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HC(hipStreamSynchronize(ctx->stream));
    size_t q = 0;
    for (int I = 0; I < ctx->nt; ++I)
        for (double& v : ctx->sp[I].params) v = p[q++];
    T.opt_host_stale = false;
    return 0;
}

int refresh_params(ace_ctx* ctx) {
    if (!ctx->params_dirty && !ctx->scale_dirty) return 0;
    std::vector<float> iw, ib;
    std::vector<double> iw64, ib64;
    if (ctx->params_dirty && upload_wpad(ctx)) return -1;
    ctx->params_dirty = false;
    if (ctx->scale_dirty) {
        for (int I = 0; I < ctx->nt; ++I) {
            const HSpecies& s = ctx->sp[I];
            for (int k = 0; k < s.K; ++k) {
                iw.push_back((float)s.inpw[k]); ib.push_back((float)s.inpb[k]);
                iw64.push_back(s.inpw[k]); ib64.push_back(s.inpb[k]);
            }
        }
        if (upload(ctx, ctx->d_inpw, iw.data(), iw.size()) || upload(ctx, ctx->d_inpb, ib.data(), ib.size()) ||
            upload(ctx, ctx->d_inpw64, iw64.data(), iw64.size()) || upload(ctx, ctx->d_inpb64, ib64.data(), ib64.size()))
            return -1;
        HC(hipStreamSynchronize(ctx->stream));   // the host vectors end here
        ctx->scale_dirty = false;
    }
    return 0;
}

int refresh_nn64(ace_ctx* ctx) {
    if (!ctx->nn64_dirty) return 0;
    if (sync_host_params(ctx)) return -1;
    std::vector<double> nn;
    for (int I = 0; I < ctx->nt; ++I) nn.insert(nn.end(), ctx->sp[I].params.begin(), ctx->sp[I].params.end());
    if (upload(ctx, ctx->d_nn64, nn.data(), nn.size())) return -1;
    HC(hipStreamSynchronize(ctx->stream));
    ctx->nn64_dirty = false;
    return 0;
}

DevBasisParams basis_params(ace_ctx* ctx) {
    DevBasisParams P;
    P.r_rs = ctx->d_r_rs.p;
    P.r_eta = ctx->d_r_eta.p;
    P.a_lamf = ctx->d_a_lamf.p;
    P.a_zeta = ctx->d_a_zeta.p;
    P.r_rs64 = ctx->d_r_rs64.p;
    P.r_eta64 = ctx->d_r_eta64.p;
    P.a_lamf64 = ctx->d_a_lamf64.p;
    P.a_zeta64 = ctx->d_a_zeta64.p;
    P.a_zint = ctx->d_a_zint.p;
    P.a_grp = ctx->d_a_grp.p;
    P.a_sel = ctx->d_a_sel.p;
    P.a_cidx = ctx->d_a_cidx.p;
    P.inpw = ctx->d_inpw.p;
    P.inpb = ctx->d_inpb.p;
    P.inpw64 = ctx->d_inpw64.p;
    P.inpb64 = ctx->d_inpb64.p;
    P.chunk_w = 3;
    return P;
}

// ---- basis kernel dispatch ------------------------------------------------------------------------

// the canonical angular accumulators are sized by the largest number of eta groups (1, 2 or 4) and of
// power levels; the generic per-function path is compiled in only when some angular block is not canonical;
// FP32 kernels of two eta groups take the model's level split
void basis_fns(ace_ctx* ctx, bool f64, BasisFn* f) {
    basis_fill(f64, ctx->ngk, ctx->nzmax, !ctx->all_canon, ctx->lsplit_on ? ctx->lsplit : 0, f);
}

int level_split(const ace_ctx* ctx, bool f64) { return (!f64 && ctx->ngk == 2 && ctx->lsplit_on) ? ctx->lsplit : 0; }

int nb_slots(ace_ctx* ctx) { return ctx->slots_cap; }

// reads the neighbour-build flags (one host round trip); on overflow the next build is synchronous
int nb_verify(ace_ctx* ctx) {
    if (!ctx->nb_pending) return 0;
    int over = 0;
    HC(hipMemcpyAsync(&over, ctx->d_scal.p + ctx->nb_over_idx, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    HC(hipStreamSynchronize(ctx->stream));
    ctx->nb_pending = false;
    if (over) {
        ctx->nb_valid = false;
        ctx->nb_async_ok = false;
        return fail(ctx, "neighbour list outgrew its capacity in an asynchronous rebuild; build it again");
    }
    return 0;
}

// CSR arrays from the key rows (export and the unbatched port); not on the energy/force path
int ensure_csr(ace_ctx* ctx) {
    if (nb_verify(ctx)) return -1;
    if (ctx->csr_valid) return 0;
    const int N = ctx->N;
    if (exclusive_scan(ctx, ctx->d_nb_count.p, ctx->d_nb_off.p, N)) return -1;
    int tot = 0;
    HC(hipMemcpyAsync(&tot, ctx->d_nb_off.p + N, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    HC(hipStreamSynchronize(ctx->stream));
    if (ensure(ctx, ctx->d_nb_j, std::max(tot, 1)) || ensure(ctx, ctx->d_nb_img, std::max(tot, 1))) return -1;
    if (timed(ctx, "nb_csr", [&] {
            k_nb_csr<<<cdiv(N, 4), 256, 0, ctx->stream>>>(N, ctx->d_struc_of.p, ctx->d_sinfo.p, ctx->d_nb_keys.p,
                                                          ctx->slots_cap, ctx->d_nb_count.p, ctx->d_nb_off.p,
                                                          ctx->d_nb_j.p, ctx->d_nb_img.p);
        }))
        return -1;
    HC(hipStreamSynchronize(ctx->stream));
    ctx->npairs = tot;
    ctx->csr_valid = true;
    return 0;
}

// mode 0: G -> out; mode 1: per-pair gradients from dE/dG; mode 2: tangent rows along eps -> out
// floats of the fused kernel's network area: staged parameters, activations and derivatives, partials
int net_lds_floats(const ace_ctx* ctx) {
    int m = 0;
    for (int I = 0; I < ctx->nt; ++I) {
        const DevSpecies& d = ctx->hdm.sp[I];
        int n = d.wp_n;
        for (int l = 0; l < d.L; ++l) n += 2 * ((d.width[l + 1] + 3) & ~3);
        m = std::max(m, n);
    }
    return m;
}

// dynamic LDS of one basis workgroup of 4 or 2 waves
int basis_bytes(ace_ctx* ctx, int mode, int waves = 4) {
    const int nbm = nb_slots(ctx), netf = mode == 3 ? net_lds_floats(ctx) : 0;
    if (waves == 2)
        return ctx->basis_f64 ? basis_lds_bytes<double, 2>(nbm, ctx->ngk, ctx->kmax, ctx->nrp, mode, netf, ctx->nt, !ctx->all_canon)
                              : basis_lds_bytes<float, 2>(nbm, ctx->ngk, ctx->kmax, ctx->nrp, mode, netf, ctx->nt, !ctx->all_canon);
    return ctx->basis_f64 ? basis_lds_bytes<double, 4>(nbm, ctx->ngk, ctx->kmax, ctx->nrp, mode, netf, ctx->nt, !ctx->all_canon)
                          : basis_lds_bytes<float, 4>(nbm, ctx->ngk, ctx->kmax, ctx->nrp, mode, netf, ctx->nt, !ctx->all_canon);
}

// The fused kernel (symmetry functions, network, adjoint of one atom per workgroup) saves two launches and
// the G / dE/dG round trips, which wins while the batch is a few rounds of resident workgroups; its LDS
// footprint limits residency, so larger batches run faster as symm -> network -> adjoint (measured on
// MI355X, 80-32-32-1 bcc: FP32 fused 0.178 / split 0.186 ms at 4000 atoms, 0.323 / 0.319 at 8000; FP64
// basis 0.128 / 0.132 at 2000, 0.230 / 0.223 at 4000).  Fused when the network fits 64 KB of LDS and the
// grid is at most four rounds of 256 CUs x the workgroups that fit in 160 KB.  Two waves per atom halve
// the per-wave gradient accumulators; where that lets more atoms share a CU the grid takes fewer rounds,
// which outweighs the longer time per atom (MI355X, 1000-atom bcc, FP64 basis: 54.6 us with two waves
// in one round against 70.3 us with four in two; FP32, one round either way: 48.6 against 43.8 us).
// the kernel configuration of path c (ACE_PATH_*; -1: the size rules) as the tuned choice of a precision
void set_path(ace_ctx* ctx, int precision, int c) {
    static const int cf[ACE_NPATHS] = {1, 0, 0, 2, 0, 0}, cs[ACE_NPATHS] = {-1, 1, 0, -1, 1, 0},
                     cw[ACE_NPATHS] = {-1, 4, 4, -1, 2, 2};
    ctx->tune_fused[precision] = c < 0 ? -1 : cf[c];
    ctx->tune_stage[precision] = c < 0 ? -1 : cs[c];
    ctx->tune_waves[precision] = c < 0 ? -1 : cw[c];
}

bool tuned(const ace_ctx* ctx, int precision) {
    return ctx->pinned[precision] || (ctx->tune_N == ctx->N && ctx->tune_slots == ctx->slots_cap);
}

long long fused_rounds(ace_ctx* ctx, int waves) {
    const long long per_round = 256LL * std::max(1, (160 * 1024) / basis_bytes(ctx, 3, waves));
    return (ctx->N + per_round - 1) / per_round;
}

// waves per atom of the fused kernel, or 0 for symm -> network -> adjoint
int fused_waves(ace_ctx* ctx, int precision) {
    if (precision == ACE_PREC_BF16L1) return 0;
    const bool fit4 = basis_bytes(ctx, 3, 4) <= 160 * 1024, fit2 = basis_bytes(ctx, 3, 2) <= 160 * 1024;
    if (tuned(ctx, precision) && ctx->tune_fused[precision] >= 0) {
        const int c = ctx->tune_fused[precision];
        return (c == 1 && fit4) ? 4 : (c == 2 && fit2) ? 2 : 0;
    }
    if (basis_bytes(ctx, 3, 4) > 64 * 1024) return 0;
    // rounds the fused kernel may take before the split path (two waves per atom) is faster; measured on
    // one 80-32-32-1 bcc structure, ms fused / split: FP32 4000 atoms (4 rounds) 0.145 / 0.150, 8000
    // 0.262 / 0.235; FP64 basis 1000 (1 round) 0.050 / 0.069, 2000 (2 rounds) 0.104 / 0.095
    const long long max_rounds = ctx->basis_f64 ? 1 : 4;
    const long long r4 = fused_rounds(ctx, 4), r2 = fused_rounds(ctx, 2);
    if (std::min(r4, r2) > max_rounds) return 0;
    return r2 < r4 ? 2 : 4;
}

// Waves per atom of the symmetry-function, adjoint and tangent kernels.  Once the batch takes several
// rounds of resident workgroups, throughput follows the atoms in flight per CU: LDS allows 160 KB / bytes
// workgroups and the wave slots (8 per SIMD) 32 / waves.  Two waves per atom keep more atoms in flight
// where LDS allows, at ~15% more time per atom (the fused kernel's measured ratio).
int split_waves(ace_ctx* ctx, int mode, int precision) {
    if (tuned(ctx, precision) && ctx->tune_waves[precision] > 0) return ctx->tune_waves[precision];
    const int b4 = basis_bytes(ctx, mode, 4), b2 = basis_bytes(ctx, mode, 2);
    const long long a4 = std::min(160 * 1024 / b4, 8), a2 = std::min(160 * 1024 / b2, 16);
    if (a4 < 1) return a2 >= 1 ? 2 : 4;
    const long long r4 = (ctx->N + 256 * a4 - 1) / (256 * a4), r2 = (ctx->N + 256 * a2 - 1) / (256 * a2);
    return (r2 * 115 < r4 * 100) ? 2 : 4;
}

// mode 0: G -> out; mode 1: per-pair gradients from dE/dG; mode 2: tangent rows along eps -> out;
// mode 3: per-atom energies and per-pair gradients in one pass.  waves: 4 or 2 per atom, -1 for
// split_waves of the precision (modes 0-2; -1: the context's precision) or four (mode 3)
int run_basis(ace_ctx* ctx, int mode, float* out = nullptr, const double* eps = nullptr, int waves = -1,
              int precision = -1, bool gather = false) {
    ++ctx->seq;
    BasisFn fns[8];
    const bool f64 = ctx->basis_f64;
    basis_fns(ctx, f64, fns);
    const int nbm = nb_slots(ctx);
    if (waves < 0) waves = mode == 3 ? 4 : split_waves(ctx, mode, precision < 0 ? ctx->precision : precision);
    if (waves != 4 && waves != 2) return fail(ctx, "basis: no kernel for this wave count");
    const int bytes = basis_bytes(ctx, mode, waves);
    if (bytes > 160 * 1024) return fail(ctx, "basis: neighbour list too long for LDS");
    BasisFn f = fns[mode + (waves == 2 ? 4 : 0)];
    if (bytes > 64 * 1024) HC(hipFuncSetAttribute((const void*)f, hipFuncAttributeMaxDynamicSharedMemorySize, bytes));
    DevBasisParams P = basis_params(ctx);
    {
        const int pr = precision < 0 ? ctx->precision : precision;
        if (tuned(ctx, pr) && ctx->tune_chunk[pr] > 0) P.chunk_w = ctx->tune_chunk[pr];
        if (ctx->pin_chunk[pr] > 0) P.chunk_w = ctx->pin_chunk[pr];
    }
    int* err = ctx->d_scal.p + 3;
    float* rows = out ? out : ctx->d_G.p;
    static const char* names[2][4] = {{"basis_symm", "basis_adjoint", "basis_tangent", "basis_fused"},
                                      {"basis_symm_f64", "basis_adjoint_f64", "basis_tangent_f64", "basis_fused_f64"}};
    BasisIO io;
    io.type = ctx->d_type.p;
    io.pos = ctx->d_pos.p;
    io.struc_of = ctx->d_struc_of.p;
    io.atom_off = ctx->d_atom_off.p;
    io.sinfo = ctx->d_sinfo.p;
    io.keys = ctx->d_nb_keys.p;
    io.counts = ctx->d_nb_count.p;
    io.cap = ctx->slots_cap;
    io.nrp = ctx->nrp;
    io.row_of = ctx->d_row_of.p;
    io.ldg = ctx->ldg;
    io.G = rows;
    io.dEdG = ctx->d_dEdG.p;
    io.err = err;
    io.lds_bytes = bytes;
    io.eps = eps;
    io.wpad = ctx->d_wpad.p;
    io.e_atom = ctx->d_e_atom.p;
    const int b = ctx->fbuf;
    io.F_fix = ctx->d_Ffix[b].p; io.E_fix = ctx->d_Efix[b].p; io.W_fix = ctx->d_Wfix[b].p;
    io.F_nxt = ctx->d_Ffix[1 - b].p; io.E_nxt = ctx->d_Efix[1 - b].p; io.W_nxt = ctx->d_Wfix[1 - b].p;
    io.gin = gather ? ctx->d_gin.p : nullptr;
    io.fown = gather ? ctx->d_fown.p : nullptr;
    io.rev = nullptr;
    const bool rev_ok = ctx->d_rev.p && ctx->d_rev.cap >= (size_t)ctx->N * ctx->slots_cap;
    if (mode == 0 && ctx->fgather && rev_ok && !f64) io.rev = ctx->d_rev.p;
    if (mode == 1 && gather && rev_ok && ctx->rev_build == ctx->nb_builds) io.rev = ctx->d_rev.p;
    if (timed(ctx, names[f64][mode], [&] {
            f<<<ctx->N, 64 * waves, bytes, ctx->stream>>>(ctx->N, ctx->d_model.p, P, io, nbm, ctx->nt, ctx->kmax);
        }))
        return -1;
    if (mode == 1 || mode == 3) ctx->fbuf = 1 - b;   // the next gradient launch adds into the zeroed buffer
    if (mode == 0 && io.rev) ctx->rev_build = ctx->nb_builds;
    return 0;
}

// the accumulators of the last gradient launch as FP64 forces, energies and virials (device)
int run_finalize(ace_ctx* ctx) {
    ++ctx->seq;
    const int b = 1 - ctx->fbuf;
    const int n = std::max(3 * ctx->N, 9 * ctx->S);
    return timed(ctx, "finalize", [&] {
        k_finalize<<<cdiv(n, 256), 256, 0, ctx->stream>>>(3 * ctx->N, ctx->S, ctx->d_Ffix[b].p, ctx->d_Efix[b].p,
                                                          ctx->d_Wfix[b].p, ctx->d_F.p, ctx->d_E.p, ctx->d_vir.p);
    });
}

// waves per 16-atom network tile: the autotuner's choice, else four
int mlp_waves(ace_ctx* ctx, int precision) {
    if (ctx->pin_mlpw[precision] > 0) return ctx->pin_mlpw[precision];
    const int w = tuned(ctx, precision) ? ctx->tune_mlpw[precision] : 0;
    return (w == 1 || w == 2 || w == 4) ? w : 4;
}

int run_mlp(ace_ctx* ctx, bool bf16, int precision = ACE_PREC_FP32) {
    ++ctx->seq;
    int lds = 0, lds_w = 0;
    for (int I = 0; I < ctx->nt; ++I) {
        lds = std::max(lds, mlp_lds_floats(ctx->hdm.sp[I]) * 4);
        lds_w = std::max(lds_w, (mlp_lds_floats(ctx->hdm.sp[I]) + ctx->hdm.sp[I].wp_n) * 4);
    }
    bool stage = lds_w <= 64 * 1024;
    if (tuned(ctx, precision) && ctx->tune_stage[precision] >= 0 && lds_w <= 160 * 1024) stage = ctx->tune_stage[precision] == 1;
    if (stage) lds = lds_w;
    if (lds > 160 * 1024) return fail(ctx, "mlp: network too wide for LDS");
    const int W = mlp_waves(ctx, precision);
#define ACE_MLP_F(W_) (stage ? (bf16 ? k_mlp_energy_grad<true, true, W_> : k_mlp_energy_grad<true, false, W_>) \
                             : (bf16 ? k_mlp_energy_grad<false, true, W_> : k_mlp_energy_grad<false, false, W_>))
    auto f = W == 1 ? ACE_MLP_F(1) : (W == 2 ? ACE_MLP_F(2) : ACE_MLP_F(4));
#undef ACE_MLP_F
    if (lds > 64 * 1024) HC(hipFuncSetAttribute((const void*)f, hipFuncAttributeMaxDynamicSharedMemorySize, lds));
    return timed(ctx, bf16 ? "mlp_energy_grad_bf16" : "mlp_energy_grad", [&] {
        f<<<(int)ctx->tiles_h.size(), 64 * W, lds, ctx->stream>>>(ctx->d_model.p, ctx->d_wpad.p,
                                                              (const __bf16*)ctx->d_w0bf.p, ctx->d_inpw.p,
                                                              ctx->d_tiles.p, ctx->ldg, ctx->d_G.p, ctx->d_dEdG.p,
                                                              ctx->d_atom_of_row.p, ctx->d_e_atom.p);
    });
}

// forces of the last gradient launch (which wrote per-pair gradients) into its fixed-point buffer
int run_force_gather(ace_ctx* ctx) {
    ++ctx->seq;
    const int b = 1 - ctx->fbuf;
    return timed(ctx, "force_gather", [&] {
        k_force_gather<<<cdiv(ctx->N, FG_WAVES), FG_WAVES * 64, 0, ctx->stream>>>(
            ctx->N, ctx->slots_cap, ctx->d_nb_count.p, ctx->d_gin.p, ctx->d_fown.p, ctx->d_Ffix[b].p);
    });
}

// the energy/force launches (fused, or symm -> network -> adjoint -> force gather); forces, energies and
// virials land in the fixed-point accumulators.  The split adjoint writes each pair's gradient once and
// the gather sums them by atom, instead of three fixed-point atomics per pair into memory.  The BF16 first
// layer runs on the matrix cores of the network kernel, so that mode takes the split path.
// the split adjoint and, with gathered forces, the gather: forces, energies and virials of the last dE/dG
// rows in the fixed-point accumulators.  The FP64-basis adjoint keeps the atomics: it holds fewer atoms
// per CU (its LDS per atom is larger), too few to hide the reverse-entry searches of its epilogue (2.5 ->
// 3.4 ms on the 108,000-atom cell with them).
int run_adjoint(ace_ctx* ctx, int precision = -1) {
    const bool gather = ctx->fgather && ctx->d_gin.p && !ctx->basis_f64;
    if (run_basis(ctx, 1, nullptr, nullptr, -1, precision, gather)) return -1;
    return gather ? run_force_gather(ctx) : 0;
}

int run_ef(ace_ctx* ctx, int precision) {
    const int fw = fused_waves(ctx, precision);
    if (fw) return run_basis(ctx, 3, nullptr, nullptr, fw);
    if (run_basis(ctx, 0, nullptr, nullptr, -1, precision)) return -1;
    if (run_mlp(ctx, precision == ACE_PREC_BF16L1, precision)) return -1;
    return run_adjoint(ctx, precision);
}

// one velocity-Verlet step: half kick with the forces of the last gradient launch, drift, energy/force
// launches at the new positions (neighbour keys of the current list, which includes the skin), half kick
int md_step(ace_ctx* ctx, int precision, double dt, int rec_cap) {
    const int N = ctx->N, nb = cdiv(N, MD_THREADS);
    if (timed(ctx, "md_kick_drift", [&] {
            k_md_kick_drift<<<nb, MD_THREADS, 0, ctx->stream>>>(N, 0.5 * dt, dt, ctx->d_type.p, ctx->d_minv.p,
                                                                ctx->d_Ffix[1 - ctx->fbuf].p, ctx->d_pos.p,
                                                                ctx->d_vel.p, ctx->d_md_ctr.p);
        }))
        return -1;
    if (run_ef(ctx, precision)) return -1;
    const int b = 1 - ctx->fbuf;
    return timed(ctx, "md_kick", [&] {
        k_md_kick<<<nb, MD_THREADS, 0, ctx->stream>>>(N, ctx->S, 0.5 * dt, ctx->d_type.p, ctx->d_minv.p,
                                                      ctx->d_minv.p + ctx->nt, ctx->d_Ffix[b].p, ctx->d_Efix[b].p,
                                                      ctx->d_vel.p, ctx->d_md_rec.p, rec_cap, ctx->d_md_ctr.p);
    });
}

// A graph of two steps leaves the accumulator buffers as it found them, so it can be replayed any number
// of times; md_exec[p] starts with buffer p.  It is captured again when a buffer was reallocated or the
// batch, precision, time step, list capacity or record size changed (all baked into its launches).
int md_graph(ace_ctx* ctx, int precision, double dt, int rec_cap, hipGraphExec_t* out) {
    const int p = ctx->fbuf;
    long long dtbits;
    memcpy(&dtbits, &dt, 8);
    const long long sig[8] = {(long long)ctx->gen, precision, dtbits, rec_cap, ctx->slots_cap, ctx->N, ctx->S, 1};
    if (ctx->md_exec[p] && memcmp(sig, ctx->md_sig[p], sizeof(sig)) == 0) {
        *out = ctx->md_exec[p];
        return 0;
    }
    if (ctx->md_exec[p]) {
        HC(hipGraphExecDestroy(ctx->md_exec[p]));
        ctx->md_exec[p] = nullptr;
    }
    const bool prof = ctx->prof;
    ctx->prof = false;
    HC(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeRelaxed));
    const int rc = (md_step(ctx, precision, dt, rec_cap) || md_step(ctx, precision, dt, rec_cap)) ? -1 : 0;
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(ctx->stream, &g);
    ctx->prof = prof;
    ctx->fbuf = p;
    if (rc || e != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return rc ? -1 : fail_hip(ctx, "md: graph capture", e);
    }
    const hipError_t ei = hipGraphInstantiate(&ctx->md_exec[p], g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ei != hipSuccess) {
        ctx->md_exec[p] = nullptr;
        return fail_hip(ctx, "md: graph instantiate", ei);
    }
    memcpy(ctx->md_sig[p], sig, sizeof(sig));
    *out = ctx->md_exec[p];
    return 0;
}

int check_err_flag(ace_ctx* ctx, const char* what) {
    int e = 0;
    HC(hipMemcpyAsync(&e, ctx->d_scal.p + 3, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    HC(hipStreamSynchronize(ctx->stream));
    if (e) return fail(ctx, std::string(what) + ": an atom has more neighbours than the LDS slots");
    return 0;
}

int require_nb(ace_ctx* ctx) {
    if (ctx->S == 0) return fail(ctx, "no structures uploaded");
    if (!ctx->nb_valid) return fail(ctx, "neighbour list not built");
    return refresh_params(ctx);
}

}  // namespace

#include "train_api.hip"

// =================================================================================================
// C-ABI
// =================================================================================================
extern "C" {

int ace_create(ace_ctx** out, const ace_model_desc* model) {
    if (!out) return -1;
    *out = nullptr;
    ace_ctx* ctx = new ace_ctx();
    *out = ctx;
    // host-side validation first: an invalid model is reported without touching the GPU
    if (parse_model(ctx, model)) return -4;
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n < 1) {
        ctx->err = "ace_create: no HIP device";
        return -2;
    }
    if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
        ctx->err = "ace_create: stream creation failed";
        return -3;
    }
    if (ensure(ctx, ctx->d_scal, 16)) return -5;
    if (build_tables(ctx)) return -6;
    if (hipGetDevice(&ctx->device) != hipSuccess) {
        ctx->err = "ace_create: no current device";
        return -2;
    }
    ctx->scratch = scratch_on(ctx->device);   // nullptr: ace_energy_forces_noalloc reports it
    return 0;
}

void ace_destroy(ace_ctx* ctx) {
    if (!ctx) return;
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    for (auto& r : ctx->pending) { (void)hipEventDestroy(r.a); (void)hipEventDestroy(r.b); }
    release(ctx->d_model); release(ctx->d_r_rs); release(ctx->d_r_eta); release(ctx->d_a_lamf);
    release(ctx->d_a_zeta); release(ctx->d_wpad); release(ctx->d_w0bf); release(ctx->d_inpw); release(ctx->d_inpb);
    release(ctx->d_a_zint); release(ctx->d_a_grp); release(ctx->d_a_sel); release(ctx->d_a_cidx); release(ctx->d_pmodel); release(ctx->d_r_rs64);
    release(ctx->d_r_eta64); release(ctx->d_a_eta64); release(ctx->d_a_zeta64); release(ctx->d_a_lamf64);
    release(ctx->d_nn64); release(ctx->d_inpw64); release(ctx->d_inpb64);
    release(ctx->d_type); release(ctx->d_struc_of); release(ctx->d_atom_off); release(ctx->d_natoms);
    release(ctx->d_row_of); release(ctx->d_atom_of_row); release(ctx->d_pos); release(ctx->d_cell);
    release(ctx->d_tiles); release(ctx->d_G); release(ctx->d_dEdG); release(ctx->d_e_atom);
    release(ctx->d_F); release(ctx->d_E); release(ctx->d_vir);
    for (int b = 0; b < 2; ++b) { release(ctx->d_Ffix[b]); release(ctx->d_Efix[b]); release(ctx->d_Wfix[b]); }
    release(ctx->d_sinfo); release(ctx->d_cell_of); release(ctx->d_wrap_of); release(ctx->d_slot_of);
    release(ctx->d_cell_start); release(ctx->d_cpos); release(ctx->d_cinfo); release(ctx->d_nb_keys);
    release(ctx->d_gin); release(ctx->d_fown); release(ctx->d_rev);
    release(ctx->d_nb_count);
    release(ctx->d_nb_off); release(ctx->d_scan_blk); release(ctx->d_scal); release(ctx->d_nb_j);
    release(ctx->d_nb_img); release(ctx->d_G64);
    release(ctx->d_vel); release(ctx->d_minv); release(ctx->d_md_rec); release(ctx->d_md_ctr);
    if (ctx->stage_done) (void)hipEventDestroy(ctx->stage_done);
    if (ctx->h_stage) (void)hipHostFree(ctx->h_stage);
    for (int q = 0; q < 2; ++q)
        if (ctx->md_exec[q]) (void)hipGraphExecDestroy(ctx->md_exec[q]);
    train_release(ctx);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

const char* ace_last_error(const ace_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int ace_input_width(const ace_ctx* ctx, int type) {
    if (!ctx || type < 0 || type >= ctx->nt) return -1;
    return ctx->sp[type].K;
}

int ace_param_count(const ace_ctx* ctx, int type) {
    if (!ctx || type < 0 || type >= ctx->nt) return -1;
    return ctx->sp[type].nparams();
}

int ace_set_params(ace_ctx* ctx, int type, const double* params, int n) {
    if (!ctx) return -1;
    if (type < 0 || type >= ctx->nt) return fail(ctx, "ace_set_params: invalid type");
    if (n != ctx->sp[type].nparams()) return fail(ctx, "ace_set_params: invalid size: vector and network mismatch");
    if (sync_host_params(ctx)) return -1;
    ctx->sp[type].params.assign(params, params + n);
    ctx->params_dirty = true;
    ++ctx->seq;
    ctx->nn64_dirty = true;
    ctx->tr.opt_dev_stale = ctx->tr.opt_on;
    return 0;
}

int ace_get_params(const ace_ctx* cctx, int type, double* params, int n) {
    // reading may refresh the host copies from the device optimiser's parameters: a cache update, the
    // parameters themselves do not change
    ace_ctx* ctx = const_cast<ace_ctx*>(cctx);
    if (!ctx || type < 0 || type >= ctx->nt) return -1;
    if (n != ctx->sp[type].nparams()) return -1;
    if (sync_host_params(ctx)) return -1;
    std::copy(ctx->sp[type].params.begin(), ctx->sp[type].params.end(), params);
    return 0;
}

int ace_set_input_scaling(ace_ctx* ctx, int type, const double* inpw, const double* inpb, int n) {
    if (!ctx) return -1;
    if (type < 0 || type >= ctx->nt) return fail(ctx, "ace_set_input_scaling: invalid type");
    if (n != ctx->sp[type].K) return fail(ctx, "ace_set_input_scaling: size differs from the input width");
    ctx->sp[type].inpw.assign(inpw, inpw + n);
    ctx->sp[type].inpb.assign(inpb, inpb + n);
    ctx->scale_dirty = true;
    ctx->tr.g_valid = false;     // cached network-input rows include the scaling
    ++ctx->seq;
    return 0;
}

int ace_upload(ace_ctx* ctx, int n_struc, const int* natoms, const double* pos, const int* type,
               const double* cell) {
    if (!ctx) return -1;
    if (n_struc < 1) return fail(ctx, "ace_upload: no structures");
    long long N = 0;
    for (int s = 0; s < n_struc; ++s) {
        if (natoms[s] < 1) return fail(ctx, "ace_upload: a structure has no atoms");
        N += natoms[s];
    }
    if (N > (1LL << 28)) return fail(ctx, "ace_upload: too many atoms");
    for (long long i = 0; i < N; ++i)
        if (type[i] < 0 || type[i] >= ctx->nt) return fail(ctx, "ace_upload: atom type out of range");
    ctx->S = n_struc;
    ctx->N = (int)N;
    ctx->natoms.assign(natoms, natoms + n_struc);
    ctx->atom_off.assign(n_struc + 1, 0);
    for (int s = 0; s < n_struc; ++s) ctx->atom_off[s + 1] = ctx->atom_off[s] + natoms[s];
    ctx->type_h.assign(type, type + N);
    ctx->cell.assign(cell, cell + 9 * n_struc);
    ctx->struc_of_h.resize(N);
    for (int s = 0; s < n_struc; ++s)
        for (int i = ctx->atom_off[s]; i < ctx->atom_off[s + 1]; ++i) ctx->struc_of_h[i] = s;
    ctx->sinfo_up.clear();
    // species-sorted rows (stable), tiles of tm rows of one species
    std::vector<int> cnt(ctx->nt, 0), start(ctx->nt + 1, 0);
    for (long long i = 0; i < N; ++i) cnt[type[i]]++;
    for (int t = 0; t < ctx->nt; ++t) start[t + 1] = start[t] + cnt[t];
    ctx->row_of_h.resize(N);
    std::vector<int> atom_of_row(N), fillp(start.begin(), start.end() - 1);
    for (long long i = 0; i < N; ++i) {
        const int r = fillp[type[i]]++;
        ctx->row_of_h[i] = r;
        atom_of_row[r] = (int)i;
    }
    // network tiles: 16 rows of one species (one matrix-core wave each)
    ctx->tm = MLP_ROWS;
    ctx->tiles_h.clear();
    for (int t = 0; t < ctx->nt; ++t)
        for (int r = start[t]; r < start[t + 1]; r += ctx->tm)
            ctx->tiles_h.push_back(make_int4(t, r, std::min(ctx->tm, start[t + 1] - r), 0));
    if (upload(ctx, ctx->d_type, type, N)) return -1;
    if (upload(ctx, ctx->d_pos, pos, 3 * N)) return -1;
    if (upload(ctx, ctx->d_struc_of, ctx->struc_of_h.data(), N)) return -1;
    if (upload(ctx, ctx->d_atom_off, ctx->atom_off.data(), n_struc + 1)) return -1;
    if (upload(ctx, ctx->d_natoms, natoms, n_struc)) return -1;
    if (upload(ctx, ctx->d_row_of, ctx->row_of_h.data(), N)) return -1;
    if (upload(ctx, ctx->d_atom_of_row, atom_of_row.data(), N)) return -1;
    if (upload(ctx, ctx->d_cell, cell, 9 * n_struc)) return -1;
    if (upload(ctx, ctx->d_tiles, ctx->tiles_h.data(), ctx->tiles_h.size())) return -1;
    if (ensure(ctx, ctx->d_G, (size_t)N * ctx->ldg)) return -1;
    if (ensure(ctx, ctx->d_dEdG, (size_t)N * ctx->ldg)) return -1;
    HC(hipMemsetAsync(ctx->d_G.p, 0, (size_t)N * ctx->ldg * sizeof(float), ctx->stream));
    HC(hipMemsetAsync(ctx->d_dEdG.p, 0, (size_t)N * ctx->ldg * sizeof(float), ctx->stream));
    if (ensure(ctx, ctx->d_e_atom, N)) return -1;
    if (ensure(ctx, ctx->d_F, 3 * N)) return -1;
    if (ensure(ctx, ctx->d_E, n_struc)) return -1;
    if (ensure(ctx, ctx->d_vir, 9 * n_struc)) return -1;
    for (int b = 0; b < 2; ++b) {
        if (ensure(ctx, ctx->d_Ffix[b], 3 * (size_t)N) || ensure(ctx, ctx->d_Efix[b], n_struc) ||
            ensure(ctx, ctx->d_Wfix[b], 9 * (size_t)n_struc))
            return -1;
        HC(hipMemsetAsync(ctx->d_Ffix[b].p, 0, 3 * (size_t)N * sizeof(unsigned long long), ctx->stream));
        HC(hipMemsetAsync(ctx->d_Efix[b].p, 0, n_struc * sizeof(unsigned long long), ctx->stream));
        HC(hipMemsetAsync(ctx->d_Wfix[b].p, 0, 9 * (size_t)n_struc * sizeof(unsigned long long), ctx->stream));
    }
    ctx->fbuf = 0;
    HC(hipStreamSynchronize(ctx->stream));
    ctx->nb_valid = false;
    ctx->nb_same_layout = false;
    ctx->nb_pending = false;
    train_on_upload(ctx);
    return 0;
}

int ace_update_positions(ace_ctx* ctx, const double* pos) {
    if (!ctx) return -1;
    if (ctx->S == 0) return fail(ctx, "ace_update_positions: no structures uploaded");
    HC(hipMemcpyAsync(ctx->d_pos.p, pos, 3 * (size_t)ctx->N * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    HC(hipStreamSynchronize(ctx->stream));
    ctx->tr.g_valid = false;
    ++ctx->seq;
    return 0;
}

int ace_build_neighbors(ace_ctx* ctx, double skin) {
    if (!ctx) return -1;
    if (ctx->S == 0) return fail(ctx, "ace_build_neighbors: no structures uploaded");
    if (skin < 0) return fail(ctx, "ace_build_neighbors: negative skin");
    const double rcl = ctx->rc + skin;
    const double rsearch = rcl * (1.0 + 1e-9) + 1e-9;
    if (rcl != ctx->nb_rc) ctx->nb_same_layout = false;
    ctx->nb_rc = rcl;
    ctx->nb_skin = skin;
    ctx->nb_valid = false;
    ctx->tr.g_valid = false;
    ++ctx->seq;
    ctx->sinfo_h.resize(ctx->S);
    int cell_off = 0;
    bool brute = true;
    for (int s = 0; s < ctx->S; ++s) {
        StrucInfo& I = ctx->sinfo_h[s];
        memset(&I, 0, sizeof(I));
        const double* R = &ctx->cell[9 * s];
        for (int q = 0; q < 9; ++q) I.R[q] = R[q];
        const double det = R[0] * (R[4] * R[8] - R[7] * R[5]) - R[3] * (R[1] * R[8] - R[7] * R[2]) +
                           R[6] * (R[1] * R[5] - R[4] * R[2]);
        double fro = 0;
        for (int q = 0; q < 9; ++q) fro += R[q] * R[q];
        if (!(std::fabs(det) > 1e-9) || !(std::sqrt(fro) > 1e-12))
            return fail(ctx, "ace_build_neighbors: singular cell (non-periodic structures are not supported)");
        mat3_inverse(R, I.Rinv);
        double hgt[3];
        for (int d = 0; d < 3; ++d) {
            const double coln = std::sqrt(R[3 * d] * R[3 * d] + R[3 * d + 1] * R[3 * d + 1] + R[3 * d + 2] * R[3 * d + 2]);
            // image range of the reference's list radius (neighbor.cpp:49-53), the model cutoff: a skin
            // widens the distance test only, so the pairs within the cutoff are the reference's
            I.cf[d] = (int)std::ceil(3.0 * ctx->rc / coln);
            // row d of R^-1 (column-major storage: (d, c) at d + 3 c)
            const double* Ri = I.Rinv;
            const double rown = std::sqrt(Ri[d] * Ri[d] + Ri[d + 3] * Ri[d + 3] + Ri[d + 6] * Ri[d + 6]);
            hgt[d] = 1.0 / rown;                                              // perpendicular height
            if (I.cf[d] > 511) return fail(ctx, "ace_build_neighbors: image range too large for this cell");
        }
        // Cells of about half the list radius: a 5-cell stencil per dimension (st = 2) scans (5/2)^3 rc^3 of
        // candidates instead of 27 rc^3 with cells of one radius.  The binning workgroup holds one counter
        // per cell, so a grid over NB_MAXC cells is coarsened; when that leaves the half-radius stencil more
        // candidate volume, prod_d (2 st_d + 1) h_d / n_d, than cells of one radius, the latter are taken.
        auto grid = [&](int div, int* nc, int* st) {
            for (int d = 0; d < 3; ++d) nc[d] = std::max(1, std::min((int)std::floor(div * hgt[d] / rsearch), 1024));
            while ((long long)nc[0] * nc[1] * nc[2] > NB_MAXC) {
                int d = 0;
                for (int e = 1; e < 3; ++e)
                    if (nc[e] > nc[d]) d = e;
                nc[d] -= 1;
            }
            double v = 1.0;
            for (int d = 0; d < 3; ++d) {
                st[d] = std::max(1, (int)std::ceil(rsearch * nc[d] / hgt[d]));
                v *= (2 * st[d] + 1) * hgt[d] / nc[d];
            }
            return v;
        };
        {
            int nc1[3], st1[3];
            const double v2 = grid(NB_CELL_DIV, I.nc, I.st), v1 = grid(1, nc1, st1);
            if (v1 < v2)
                for (int d = 0; d < 3; ++d) {
                    I.nc[d] = nc1[d];
                    I.st[d] = st1[d];
                }
        }
        // every structure small and with heights over twice the list radius: search without binning
        if (ctx->natoms[s] > NB_BRUTE_MAX || !(hgt[0] > 2 * rsearch && hgt[1] > 2 * rsearch && hgt[2] > 2 * rsearch))
            brute = false;
        const int ncell = I.nc[0] * I.nc[1] * I.nc[2];
        I.atom_off = ctx->atom_off[s];
        I.natoms = ctx->natoms[s];
        I.cell_off = cell_off;
        if ((long long)cell_off + ncell + ctx->S > (1LL << 30)) return fail(ctx, "ace_build_neighbors: too many cells");
        cell_off += ncell;
    }
    ctx->ncells = cell_off;
    const int N = ctx->N, ncs = ctx->ncells + ctx->S;
    // the cell geometry changes only with the cells or the cutoff: skip the upload when it is unchanged
    if (ctx->sinfo_up.size() != ctx->sinfo_h.size() ||
        memcmp(ctx->sinfo_up.data(), ctx->sinfo_h.data(), ctx->S * sizeof(StrucInfo)) != 0 || !ctx->d_sinfo.p) {
        if (upload(ctx, ctx->d_sinfo, ctx->sinfo_h.data(), ctx->S)) return -1;
        ctx->sinfo_up = ctx->sinfo_h;
    }
    if (ensure(ctx, ctx->d_cell_of, N) || ensure(ctx, ctx->d_wrap_of, N) || ensure(ctx, ctx->d_slot_of, N) ||
        ensure(ctx, ctx->d_cell_start, ncs) || ensure(ctx, ctx->d_cpos, N) || ensure(ctx, ctx->d_cinfo, N) ||
        ensure(ctx, ctx->d_nb_count, N + 1) || ensure(ctx, ctx->d_nb_off, N + 1))
        return -1;
    if (!brute && timed(ctx, "nb_bin", [&] {
            k_nb_bin<<<ctx->S, NB_BIN_THREADS, 0, ctx->stream>>>(ctx->d_pos.p, ctx->d_sinfo.p, ctx->d_cell_of.p,
                                                                 ctx->d_wrap_of.p, ctx->d_slot_of.p,
                                                                 ctx->d_cell_start.p, ctx->d_cpos.p, ctx->d_cinfo.p,
                                                                 ctx->d_scal.p);
        }))
        return -1;
    ctx->nb_brute = brute;
    // the brute-force build's overflow flag alternates between d_scal[6] and d_scal[7]: each build clears
    // the other one for the next (the binning kernel clears all flags otherwise)
    const int p_over = 6 + ctx->nb_parity, p_next = 7 - ctx->nb_parity;
    const double rc2 = rcl * rcl;
    // Capacity: the first build (or one after an overflow) counts first and reads the largest count back;
    // later builds reuse the key-row capacity (the LDS slots of the basis kernels, with a margin) and run
    // without a host round trip.  An atom that outgrows it raises a flag checked at the next synchronising
    // call, and the basis kernels refuse it.
    const bool async = ctx->nb_async_ok && ctx->nb_same_layout;
    if (!async) {
        if (brute) HC(hipMemsetAsync(ctx->d_scal.p, 0, 16 * sizeof(int), ctx->stream));
        if (timed(ctx, "nb_count", [&] {
                if (brute)
                    k_nb_brute<0><<<N, NB_WAVES * 64, 0, ctx->stream>>>(
                        N, rc2, ctx->d_pos.p, ctx->d_struc_of.p, ctx->d_sinfo.p, ctx->d_nb_count.p, ctx->d_scal.p,
                        nullptr, 0, ctx->d_scal.p + p_over, ctx->d_scal.p + p_next, ctx->d_scal.p + 3);
                else
                    k_nb_search<0><<<N, NB_SEARCH_WAVES * 64, 0, ctx->stream>>>(
                        N, rc2, ctx->d_pos.p, ctx->d_struc_of.p, ctx->d_sinfo.p, ctx->d_cell_of.p, ctx->d_wrap_of.p,
                        ctx->d_cell_start.p, ctx->d_cpos.p, ctx->d_cinfo.p, ctx->d_nb_count.p, ctx->d_scal.p, nullptr,
                        0, nullptr);
            }))
            return -1;
        int mx = 0;
        HC(hipMemcpyAsync(&mx, ctx->d_scal.p, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
        HC(hipStreamSynchronize(ctx->stream));
        if (mx > NB_LMAX) return fail(ctx, "ace_build_neighbors: an atom has more than 1024 neighbours");
        // the largest count plus a margin of 8 for later asynchronous rebuilds, rounded up to 8: the basis
        // kernels' LDS is sized by it, and every slot less keeps more atoms in flight per CU (the 568k-atom
        // batch's adjoint holds 11 atoms per CU at 136 slots, 10 at 144); a row that outgrows it costs a
        // counting rebuild
        ctx->slots_cap = std::min(NB_LMAX, ((mx + 8 + 7) & ~7));
    }
    if (ensure(ctx, ctx->d_nb_keys, (size_t)N * ctx->slots_cap)) return -1;
    if (ctx->fgather && (ensure(ctx, ctx->d_gin, 3 * (size_t)N * ctx->slots_cap) || ensure(ctx, ctx->d_fown, 3 * (size_t)N)))
        return -1;
    if (ctx->fgather && (long long)N * ctx->slots_cap < (1LL << 31) && ensure(ctx, ctx->d_rev, (size_t)N * ctx->slots_cap))
        return -1;
    if (timed(ctx, "nb_search", [&] {
            if (brute)
                k_nb_brute<1><<<N, NB_WAVES * 64, ctx->slots_cap * sizeof(long long), ctx->stream>>>(
                    N, rc2, ctx->d_pos.p, ctx->d_struc_of.p, ctx->d_sinfo.p, ctx->d_nb_count.p, ctx->d_scal.p,
                    ctx->d_nb_keys.p, ctx->slots_cap, ctx->d_scal.p + p_over, ctx->d_scal.p + p_next,
                    ctx->d_scal.p + 3);
            else
                k_nb_search<1><<<N, NB_SEARCH_WAVES * 64, ctx->slots_cap * sizeof(long long), ctx->stream>>>(
                    N, rc2, ctx->d_pos.p, ctx->d_struc_of.p, ctx->d_sinfo.p, ctx->d_cell_of.p, ctx->d_wrap_of.p,
                    ctx->d_cell_start.p, ctx->d_cpos.p, ctx->d_cinfo.p, ctx->d_nb_count.p, ctx->d_scal.p,
                    ctx->d_nb_keys.p, ctx->slots_cap, ctx->d_scal.p + 5);
        }))
        return -1;
    ctx->nb_over_idx = brute ? p_over : 5;
    if (brute) ctx->nb_parity ^= 1;
    ctx->nb_valid = true;
    ++ctx->nb_builds;
    ctx->nb_pending = true;
    ctx->csr_valid = false;
    if (!async) {
        if (nb_verify(ctx)) return -1;
        ctx->nb_async_ok = true;
    }
    ctx->nb_same_layout = true;
    return flush_prof(ctx);
}

int ace_neighbor_count(const ace_ctx* cctx, long long* npairs) {
    ace_ctx* ctx = const_cast<ace_ctx*>(cctx);
    if (!ctx || !npairs || !ctx->nb_valid) return -1;
    if (ensure_csr(ctx)) return -1;
    *npairs = ctx->npairs;
    return 0;
}

int ace_export_neighbors(ace_ctx* ctx, long long* offsets, int* j, int* img) {
    if (!ctx) return -1;
    if (!ctx->nb_valid) return fail(ctx, "ace_export_neighbors: neighbour list not built");
    if (ensure_csr(ctx)) return -1;
    const int N = ctx->N;
    const long long P = ctx->npairs;
    std::vector<int> off(N + 1), jj(P), im(P);
    HC(hipMemcpyAsync(off.data(), ctx->d_nb_off.p, (N + 1) * sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    if (P) {
        HC(hipMemcpyAsync(jj.data(), ctx->d_nb_j.p, P * sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
        HC(hipMemcpyAsync(im.data(), ctx->d_nb_img.p, P * sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    }
    HC(hipStreamSynchronize(ctx->stream));
    for (int i = 0; i <= N; ++i) offsets[i] = off[i];
    for (int i = 0; i < N; ++i) {
        const int a0 = ctx->atom_off[ctx->struc_of_h[i]];
        for (int p = off[i]; p < off[i + 1]; ++p) {
            j[p] = jj[p] - a0;
            img_unpack(im[p], img[3 * p], img[3 * p + 1], img[3 * p + 2]);
        }
    }
    return 0;
}

int ace_set_precision(ace_ctx* ctx, int precision) {
    if (!ctx) return -1;
    if (precision != ACE_PREC_FP32 && precision != ACE_PREC_FP64_BASIS)
        return fail(ctx, "ace_set_precision: precision mode not available");
    if (precision != ctx->precision) ctx->tr.g_valid = false;
    ++ctx->seq;
    ctx->precision = precision;
    return 0;
}

int ace_symm(ace_ctx* ctx, double* G, int ldg) {
    if (!ctx) return -1;
    ctx->basis_f64 = ctx->precision == ACE_PREC_FP64_BASIS;
    if (require_nb(ctx)) return -1;
    HC(hipMemsetAsync(ctx->d_scal.p + 3, 0, sizeof(int), ctx->stream));
    if (run_basis(ctx, 0)) return -1;
    if (check_err_flag(ctx, "ace_symm")) return -1;
    if (ctx->nb_pending && nb_verify(ctx)) return -1;
    if (G) {
        if (ldg < ctx->kmax) return fail(ctx, "ace_symm: ldg smaller than the widest input vector");
        DBuf<double> tmp;
        if (ensure(ctx, tmp, (size_t)ctx->N * ldg)) return -1;
        dim3 grid(cdiv(ldg, 128), ctx->N);
        k_export_rows<<<grid, 128, 0, ctx->stream>>>(ctx->N, ctx->d_row_of.p, ctx->d_type.p, ctx->d_model.p,
                                                    ctx->d_inpw64.p, ctx->d_inpb64.p, ctx->ldg, ctx->d_G.p, ldg,
                                                    tmp.p);
        hipError_t e = hipGetLastError();
        if (e == hipSuccess) e = hipMemcpyAsync(G, tmp.p, (size_t)ctx->N * ldg * sizeof(double), hipMemcpyDeviceToHost, ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        release(tmp);
        if (e != hipSuccess) return fail_hip(ctx, "ace_symm export", e);
    }
    return flush_prof(ctx);
}

int ace_energy_forces(ace_ctx* ctx, int precision, double* E_struc, double* F, double* virial) {
    if (!ctx) return -1;
    if (precision != ACE_PREC_FP32 && precision != ACE_PREC_FP64_BASIS && precision != ACE_PREC_BF16L1)
        return fail(ctx, "ace_energy_forces: precision mode not available");
    ctx->basis_f64 = precision == ACE_PREC_FP64_BASIS;
    if (require_nb(ctx)) return -1;
    if (run_ef(ctx, precision)) return -1;
    if (E_struc || F || virial || ctx->prof) {
        if (ctx->nb_pending && nb_verify(ctx)) {
            // an asynchronous rebuild outgrew its capacity: rebuild with the exact size and recompute
            if (ctx->nb_async_ok) return -1;
            if (ace_build_neighbors(ctx, ctx->nb_skin)) return -1;
            return ace_energy_forces(ctx, precision, E_struc, F, virial);
        }
        if (check_err_flag(ctx, "ace_energy_forces")) return -1;
    }
    if (E_struc || F || virial) {
        if (run_finalize(ctx)) return -1;
        if (E_struc) HC(hipMemcpyAsync(E_struc, ctx->d_E.p, ctx->S * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        if (F) HC(hipMemcpyAsync(F, ctx->d_F.p, 3 * (size_t)ctx->N * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        if (virial) HC(hipMemcpyAsync(virial, ctx->d_vir.p, 9 * ctx->S * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        HC(hipStreamSynchronize(ctx->stream));
    }
    return flush_prof(ctx);
}

int ace_compute(ace_ctx* ctx, int precision, double skin, double* E_struc, double* F, double* virial) {
    if (!ctx) return -1;
    if (ace_build_neighbors(ctx, skin)) return -1;
    return ace_energy_forces(ctx, precision, E_struc, F, virial);
}

int ace_energy_forces_noalloc(ace_ctx* ctx, int precision, const double* pos, double* E_struc, double* F,
                              double* virial) {
    if (!ctx) return -1;
    if (precision != ACE_PREC_FP32 && precision != ACE_PREC_FP64_BASIS && precision != ACE_PREC_BF16L1)
        return fail(ctx, "ace_energy_forces_noalloc: precision mode not available");
    if (!ctx->scratch) return fail(ctx, "ace_energy_forces_noalloc: the scratch array of this device was not found");
    int dev = -1;
    HC(hipGetDevice(&dev));
    if (dev != ctx->device) return fail(ctx, "ace_energy_forces_noalloc: call it with the context's device current");
    if (!E_struc || !F) return fail(ctx, "ace_energy_forces_noalloc: E_struc and F (device pointers) are required");
    if (ctx->S == 0) return fail(ctx, "ace_energy_forces_noalloc: no structures uploaded");
    if (!ctx->nb_valid) return fail(ctx, "ace_energy_forces_noalloc: neighbour list not built");
    if (ctx->params_dirty || ctx->scale_dirty)
        return fail(ctx, "ace_energy_forces_noalloc: parameters or input scaling changed since the last "
                         "ace_energy_forces / ace_symm call, which uploads them (this entry point does not)");
    ctx->basis_f64 = precision == ACE_PREC_FP64_BASIS;
    const int N = ctx->N, S = ctx->S;
    const bool split = fused_waves(ctx, precision) == 0;
    // scratch layout: two accumulator sets (forces 3N, energies S, virials 9S), per-atom energies, the network
    // rows G and dE/dG of the split path, and the virial when the caller passes none
    auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
    const size_t acc_w = 3 * (size_t)N + 10 * (size_t)S;
    const size_t o_acc1 = al(8 * acc_w), o_e = o_acc1 + al(8 * acc_w), o_g = o_e + al(8 * (size_t)N);
    const size_t rows = split ? al(4 * (size_t)N * ctx->ldg) : 0;
    const size_t o_d = o_g + rows, o_w = o_d + rows;
    const size_t need = o_w + (virial ? 0 : al(8 * 9 * (size_t)S));
    if (need > ACE_NOALLOC_BYTES) {
        char m[256];
        snprintf(m, sizeof m, "ace_energy_forces_noalloc: this batch needs %zu bytes of scratch, the maximum is %llu",
                 need, (unsigned long long)ACE_NOALLOC_BYTES);
        return fail(ctx, m);
    }
    if (pos) {
        HC(hipMemcpyAsync(ctx->d_pos.p, pos, 3 * (size_t)N * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
        ctx->tr.g_valid = false;
    ++ctx->seq;
    }
    unsigned char* sp = ctx->scratch;
    auto u64 = [&](size_t o) { return (unsigned long long*)(sp + o); };
    // the context's buffers, pointed into scratch and the caller's outputs for this call
    DBuf<unsigned long long> sF[2] = {ctx->d_Ffix[0], ctx->d_Ffix[1]}, sE[2] = {ctx->d_Efix[0], ctx->d_Efix[1]},
                             sW[2] = {ctx->d_Wfix[0], ctx->d_Wfix[1]};
    const DBuf<float> sG = ctx->d_G, sD = ctx->d_dEdG;
    const DBuf<double> se = ctx->d_e_atom, sFo = ctx->d_F, sEo = ctx->d_E, svo = ctx->d_vir;
    const int sfbuf = ctx->fbuf;
    const bool sprof = ctx->prof;
    for (int b = 0; b < 2; ++b) {
        const size_t o = b ? o_acc1 : 0;
        ctx->d_Ffix[b].p = u64(o);
        ctx->d_Efix[b].p = u64(o) + 3 * (size_t)N;
        ctx->d_Wfix[b].p = u64(o) + 3 * (size_t)N + S;
    }
    ctx->d_e_atom.p = (double*)(sp + o_e);
    if (split) {
        ctx->d_G.p = (float*)(sp + o_g);
        ctx->d_dEdG.p = (float*)(sp + o_d);
    }
    ctx->d_F.p = F;
    ctx->d_E.p = E_struc;
    ctx->d_vir.p = virial ? virial : (double*)(sp + o_w);
    ctx->fbuf = 0;
    ctx->prof = false;
    // the array keeps the last call's contents: the accumulators this call adds into start at zero, and the
    // network rows (whose padding columns the matrix-core tiles read) too
    auto zero = [&](size_t o, size_t words) {
        const int grid = (int)std::min<size_t>((words + 255) / 256, 4096);
        return timed(ctx, "zero", [&] { k_zero_u64<<<grid, 256, 0, ctx->stream>>>(words, u64(o)); });
    };
    int rc = zero(0, acc_w);
    if (!rc && split) rc = zero(o_g, (o_w - o_g) / 8);
    if (!rc) rc = run_ef(ctx, precision);
    if (!rc) rc = run_finalize(ctx);
    for (int b = 0; b < 2; ++b) {
        ctx->d_Ffix[b] = sF[b];
        ctx->d_Efix[b] = sE[b];
        ctx->d_Wfix[b] = sW[b];
    }
    ctx->d_G = sG;
    ctx->d_dEdG = sD;
    ctx->d_e_atom = se;
    ctx->d_F = sFo;
    ctx->d_E = sEo;
    ctx->d_vir = svo;
    ctx->fbuf = sfbuf;
    ctx->prof = sprof;
    return rc;
}

int ace_md_init(ace_ctx* ctx, const double* mass, const double* vel) {
    if (!ctx) return -1;
    if (ctx->S == 0) return fail(ctx, "ace_md_init: no structures uploaded");
    if (!mass) return fail(ctx, "ace_md_init: masses required");
    std::vector<double> mi(2 * ctx->nt);
    for (int I = 0; I < ctx->nt; ++I) {
        if (!(mass[I] > 0)) return fail(ctx, "ace_md_init: masses must be positive");
        mi[I] = 1.0 / (mass[I] * MD_MVV2E);
        mi[ctx->nt + I] = mass[I] * MD_MVV2E;
    }
    if (upload(ctx, ctx->d_minv, mi.data(), mi.size())) return -1;
    if (ensure(ctx, ctx->d_vel, 3 * (size_t)ctx->N) || ensure(ctx, ctx->d_md_ctr, 2)) return -1;
    if (vel)
        HC(hipMemcpyAsync(ctx->d_vel.p, vel, 3 * (size_t)ctx->N * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    else
        HC(hipMemsetAsync(ctx->d_vel.p, 0, 3 * (size_t)ctx->N * sizeof(double), ctx->stream));
    HC(hipStreamSynchronize(ctx->stream));
    ctx->md_ready = true;
    return 0;
}

int ace_md_run(ace_ctx* ctx, int precision, int nsteps, double dt, int rebuild_every, double skin, int use_graph,
               double* energy) {
    if (!ctx) return -1;
    if (!ctx->md_ready) return fail(ctx, "ace_md_run: call ace_md_init first");
    if (precision != ACE_PREC_FP32 && precision != ACE_PREC_FP64_BASIS && precision != ACE_PREC_BF16L1)
        return fail(ctx, "ace_md_run: precision mode not available");
    if (nsteps < 0 || !(dt > 0) || rebuild_every < 1 || !(skin >= 0))
        return fail(ctx, "ace_md_run: need nsteps >= 0, dt > 0, rebuild_every >= 1, skin >= 0");
    ctx->basis_f64 = precision == ACE_PREC_FP64_BASIS;
    const int rec_cap = std::max(nsteps, 1);
    if (ensure(ctx, ctx->d_md_rec, 2 * (size_t)rec_cap)) return -1;
    // record and step counter cleared outside any captured graph
    HC(hipMemsetAsync(ctx->d_md_rec.p, 0, 2 * (size_t)rec_cap * sizeof(unsigned long long), ctx->stream));
    HC(hipMemsetAsync(ctx->d_md_ctr.p, 0, 2 * sizeof(unsigned int), ctx->stream));   // step counter, flags
    auto collect_flags = [&] {
        return timed(ctx, "md_flags", [&] { k_md_flags<<<1, 64, 0, ctx->stream>>>(ctx->d_scal.p, ctx->nb_over_idx, ctx->d_md_ctr.p + 1); });
    };
    // A rebuild that reuses the row capacity is checked before any step uses it (one host round trip per
    // rebuild); if an atom outgrew the rows, the list is built again with a counting pass and resized rows
    // (the graphs are then captured again).  The flags left by the previous list are collected first, except
    // before the run's first build (they belong to an earlier call).
    auto rebuild = [&](bool collect) -> int {
        if (collect && collect_flags()) return -1;
        if (timed(ctx, "md_wrap", [&] {
                k_md_wrap<<<cdiv(ctx->N, MD_THREADS), MD_THREADS, 0, ctx->stream>>>(ctx->N, ctx->d_struc_of.p,
                                                                                  ctx->d_cell.p, ctx->d_pos.p);
            }))
            return -1;
        if (ace_build_neighbors(ctx, skin)) return -1;
        if (!ctx->nb_pending) return 0;
        int over = 0;
        HC(hipMemcpyAsync(&over, ctx->d_scal.p + ctx->nb_over_idx, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
        HC(hipStreamSynchronize(ctx->stream));
        ctx->nb_pending = false;
        if (!over) return 0;
        ctx->nb_async_ok = false;
        return ace_build_neighbors(ctx, skin);
    };
    // list radius rc + skin; forces at the starting positions
    if (rebuild(false)) return -1;
    if (require_nb(ctx)) return -1;
    if (run_ef(ctx, precision)) return -1;
    ctx->tr.g_valid = false;
    ++ctx->seq;
    for (int k = 0; k < nsteps;) {
        if (k > 0 && k % rebuild_every == 0 && rebuild(true)) return -1;
        const int c = std::min(nsteps - k, rebuild_every - k % rebuild_every);
        int done = 0;
        if (use_graph) {
            for (; done + 2 <= c; done += 2) {
                hipGraphExec_t ex = nullptr;
                if (md_graph(ctx, precision, dt, rec_cap, &ex)) return -1;
                HC(hipGraphLaunch(ex, ctx->stream));
            }
        }
        for (; done < c; ++done)
            if (md_step(ctx, precision, dt, rec_cap)) return -1;
        k += c;
    }
    if (collect_flags()) return -1;
    unsigned int bad = 0;
    HC(hipMemcpyAsync(&bad, ctx->d_md_ctr.p + 1, sizeof(bad), hipMemcpyDeviceToHost, ctx->stream));
    HC(hipStreamSynchronize(ctx->stream));
    if (ctx->nb_pending && nb_verify(ctx)) return -1;
    if (bad) {
        ctx->nb_async_ok = false;   // the next build counts again and sizes the rows to the new maximum
        return fail(ctx, "ace_md_run: an atom outgrew the neighbour-list capacity during the run; run again");
    }
    if (energy && nsteps > 0) {
        std::vector<long long> r(2 * (size_t)nsteps);
        HC(hipMemcpyAsync(r.data(), ctx->d_md_rec.p, r.size() * sizeof(long long), hipMemcpyDeviceToHost, ctx->stream));
        HC(hipStreamSynchronize(ctx->stream));
        for (size_t q = 0; q < r.size(); ++q) energy[q] = (double)r[q] * MD_FIX_INV;
    }
    return flush_prof(ctx);
}

int ace_md_state(ace_ctx* ctx, double* pos, double* vel) {
    if (!ctx) return -1;
    if (!ctx->md_ready) return fail(ctx, "ace_md_state: call ace_md_init first");
    if (pos) HC(hipMemcpyAsync(pos, ctx->d_pos.p, 3 * (size_t)ctx->N * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    if (vel) HC(hipMemcpyAsync(vel, ctx->d_vel.p, 3 * (size_t)ctx->N * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HC(hipStreamSynchronize(ctx->stream));
    if (pos) {
        // positions in their cells, as the reference's Verlet keeps them (Cell::returnToCell after each drift);
        // on the device they are returned to the cells at each list rebuild
        for (int i = 0; i < ctx->N; ++i) {
            const double* R = &ctx->cell[9 * (size_t)ctx->struc_of_h[i]];
            double Ri[9];
            mat3_inverse(R, Ri);
            double* r = pos + 3 * (size_t)i;
            double f[3];
            for (int a = 0; a < 3; ++a) f[a] = Ri[a] * r[0] + Ri[3 + a] * r[1] + Ri[6 + a] * r[2];
            for (int a = 0; a < 3; ++a) f[a] -= std::floor(f[a]);
            for (int a = 0; a < 3; ++a) r[a] = R[a] * f[0] + R[3 + a] * f[1] + R[6 + a] * f[2];
        }
    }
    return 0;
}

int ace_autotune(ace_ctx* ctx, int precision, double* ms) {
    if (!ctx) return -1;
    if (precision != ACE_PREC_FP32 && precision != ACE_PREC_FP64_BASIS && precision != ACE_PREC_BF16L1)
        return fail(ctx, "ace_autotune: precision mode not available");
    ctx->basis_f64 = precision == ACE_PREC_FP64_BASIS;
    if (require_nb(ctx)) return -1;
    if (ctx->nb_pending && nb_verify(ctx)) return -1;
    int lds_w = 0;
    for (int I = 0; I < ctx->nt; ++I) lds_w = std::max(lds_w, (mlp_lds_floats(ctx->hdm.sp[I]) + ctx->hdm.sp[I].wp_n) * 4);
    // candidates (ACE_PATH_*): fused with four waves per atom; split with staged weights; split with weights
    // read from memory; fused with two waves per atom; the two split forms with two waves per atom
    const bool fu = precision != ACE_PREC_BF16L1;
    const bool can_fused = fu && basis_bytes(ctx, 3, 4) <= 160 * 1024;
    const bool can_fused2 = fu && basis_bytes(ctx, 3, 2) <= 160 * 1024;
    const bool can_stage = lds_w <= 160 * 1024;
    const bool ok[ACE_NPATHS] = {can_fused, can_stage, true, can_fused2, can_stage, true};
    double t[ACE_NPATHS] = {-1, -1, -1, -1, -1, -1};
    const bool prof = ctx->prof;
    ctx->prof = false;
    ctx->tune_N = ctx->N;
    ctx->tune_slots = ctx->slots_cap;
    hipEvent_t e0, e1;
    HC(hipEventCreate(&e0));
    HC(hipEventCreate(&e1));
    // milliseconds of one energy/force pass in the current setting (a warm-up, then 3 or 20 timed passes)
    auto time_ef = [&](double& ms_out) -> int {
        if (run_ef(ctx, precision)) return -1;   // warm-up (kernel loads, LDS attributes)
        float el = 0;
        int reps = 3;
        for (int pass = 0; pass < 2; ++pass) {
            HC(hipEventRecord(e0, ctx->stream));
            for (int r = 0; r < reps; ++r)
                if (run_ef(ctx, precision)) return -1;
            HC(hipEventRecord(e1, ctx->stream));
            HC(hipEventSynchronize(e1));
            HC(hipEventElapsedTime(&el, e0, e1));
            if (el / reps > 0.5f) break;
            reps = 20;                               // short launches: time more of them
        }
        ms_out = (double)el / reps;
        return 0;
    };
    ctx->tune_chunk[precision] = -1;
    for (int k = 0; k < 3; ++k) ctx->tune_chunk_ms[precision][k] = -1;
    ctx->tune_mlpw[precision] = 0;
    int best = -1;
    for (int c = 0; c < ACE_NPATHS; ++c) {
        if (!ok[c]) continue;
        set_path(ctx, precision, c);
        if (time_ef(t[c])) return -1;
        if (best < 0 || t[c] < t[best]) best = c;
    }
    // the angular tile shape on the chosen path: the chunk weight w of the triplet loops
    set_path(ctx, precision, best);
    ctx->pinned[precision] = false;
    static const int wts[3] = {1, 3, 8};
    int wbest = -1;
    for (int k = 0; k < 3; ++k) {
        ctx->tune_chunk[precision] = wts[k];
        if (time_ef(ctx->tune_chunk_ms[precision][k])) return -1;
        if (wbest < 0 || ctx->tune_chunk_ms[precision][k] < ctx->tune_chunk_ms[precision][wbest]) wbest = k;
    }
    ctx->tune_chunk[precision] = wts[wbest];
    // the network tile shape on a split path: waves per 16-atom tile
    ctx->tune_mlpw[precision] = 0;
    for (int k = 0; k < 3; ++k) ctx->tune_mlpw_ms[precision][k] = -1;
    if (!fused_waves(ctx, precision)) {
        static const int nws[3] = {1, 2, 4};
        int nbest = -1;
        for (int k = 0; k < 3; ++k) {
            ctx->tune_mlpw[precision] = nws[k];
            if (time_ef(ctx->tune_mlpw_ms[precision][k])) return -1;
            if (nbest < 0 || ctx->tune_mlpw_ms[precision][k] < ctx->tune_mlpw_ms[precision][nbest]) nbest = k;
        }
        ctx->tune_mlpw[precision] = nws[nbest];
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    ctx->prof = prof;
    if (check_err_flag(ctx, "ace_autotune")) return -1;
    if (ms) {
        for (int c = 0; c < ACE_NPATHS; ++c) ms[c] = t[c];
        ms[ACE_NPATHS] = best;
    }
    return 0;
}

int ace_tile_shape(const ace_ctx* cctx, int precision, double* shape) {
    if (!cctx) return -1;
    ace_ctx* ctx = const_cast<ace_ctx*>(cctx);
    if (precision != ACE_PREC_FP32 && precision != ACE_PREC_FP64_BASIS && precision != ACE_PREC_BF16L1)
        return fail(ctx, "ace_tile_shape: precision mode not available");
    if (!shape) return fail(ctx, "ace_tile_shape: null array");
    const bool tu = tuned(ctx, precision);
    int path = -1;
    if (tu && ctx->tune_fused[precision] >= 0) {
        const int f = ctx->tune_fused[precision], w = ctx->tune_waves[precision], st = ctx->tune_stage[precision];
        path = f == 1 ? ACE_PATH_FUSED : f == 2 ? ACE_PATH_FUSED_2W
             : (w == 2 ? (st == 1 ? ACE_PATH_SPLIT_STAGED_2W : ACE_PATH_SPLIT_2W)
                       : (st == 1 ? ACE_PATH_SPLIT_STAGED : ACE_PATH_SPLIT));
    }
    const int fw = fused_waves(ctx, precision);
    shape[0] = path;
    shape[1] = fw ? fw : split_waves(ctx, 0, precision);
    shape[2] = (tu && ctx->tune_stage[precision] >= 0) ? ctx->tune_stage[precision] : -1;
    const int w = ctx->pin_chunk[precision] > 0 ? ctx->pin_chunk[precision]
                : ((tu && ctx->tune_chunk[precision] > 0) ? ctx->tune_chunk[precision] : 3);
    shape[3] = w;
    const int kw = w == 1 ? 0 : (w == 3 ? 1 : 2);
    shape[4] = tu ? ctx->tune_chunk_ms[precision][kw] : -1;
    for (int k = 0; k < 3; ++k) shape[5 + k] = tu ? ctx->tune_chunk_ms[precision][k] : -1;
    shape[8] = level_split(ctx, precision == ACE_PREC_FP64_BASIS);
    shape[9] = fw ? 0 : mlp_waves(ctx, precision);
    for (int k = 0; k < 3; ++k) shape[10 + k] = tu ? ctx->tune_mlpw_ms[precision][k] : -1;
    return 0;
}

int ace_set_tile_shape(ace_ctx* ctx, int precision, int chunk_w, int net_waves) {
    if (!ctx) return -1;
    if (precision != ACE_PREC_FP32 && precision != ACE_PREC_FP64_BASIS && precision != ACE_PREC_BF16L1)
        return fail(ctx, "ace_set_tile_shape: precision mode not available");
    if (chunk_w != 0 && chunk_w != 1 && chunk_w != 3 && chunk_w != 8)
        return fail(ctx, "ace_set_tile_shape: chunk weight must be 1, 3, 8 or 0 (tuned or default)");
    if (net_waves != 0 && net_waves != 1 && net_waves != 2 && net_waves != 4)
        return fail(ctx, "ace_set_tile_shape: waves per network tile must be 1, 2, 4 or 0 (tuned or default)");
    ctx->pin_chunk[precision] = chunk_w;
    ctx->pin_mlpw[precision] = net_waves;
    return 0;
}

int ace_set_level_split(ace_ctx* ctx, int enable) {
    if (!ctx) return -1;
    ctx->lsplit_on = enable != 0;
    return 0;
}

int ace_set_force_gather(ace_ctx* ctx, int enable) {
    if (!ctx) return -1;
    ctx->fgather = enable != 0;
    if (!ctx->fgather) {
        release(ctx->d_gin);
        release(ctx->d_fown);
        release(ctx->d_rev);
        ++ctx->gen;   // graphs captured with the gather are stale
    }
    return 0;
}

int ace_set_path(ace_ctx* ctx, int precision, int path) {
    if (!ctx) return -1;
    if (precision != ACE_PREC_FP32 && precision != ACE_PREC_FP64_BASIS && precision != ACE_PREC_BF16L1)
        return fail(ctx, "ace_set_path: precision mode not available");
    if (path < -1 || path >= ACE_NPATHS) return fail(ctx, "ace_set_path: path must be -1 (size rule) or an ACE_PATH_*");
    if (precision == ACE_PREC_BF16L1 && (path == ACE_PATH_FUSED || path == ACE_PATH_FUSED_2W))
        return fail(ctx, "ace_set_path: the BF16 first layer runs on the split path only");
    set_path(ctx, precision, path);
    ctx->pinned[precision] = path >= 0;
    return 0;
}

int ace_plain_energy_forces(ace_ctx* ctx, double* E_struc, double* F) {
    if (!ctx) return -1;
    if (!ctx->plain_ok) return fail(ctx, "ace_plain_energy_forces: model outside the port's forms (" + ctx->plain_why + ")");
    if (require_nb(ctx) || refresh_nn64(ctx)) return -1;
    if (ensure_csr(ctx)) return -1;
    int nmax = 0;
    for (int s = 0; s < ctx->S; ++s) nmax = std::max(nmax, ctx->natoms[s]);
    if (ensure(ctx, ctx->d_G64, (size_t)nmax * PL_KMAX)) return -1;
    PlainParams P;
    P.r_rs = ctx->d_r_rs64.p; P.r_eta = ctx->d_r_eta64.p; P.a_eta = ctx->d_a_eta64.p; P.a_zeta = ctx->d_a_zeta64.p;
    P.a_lamf = ctx->d_a_lamf64.p; P.nn = ctx->d_nn64.p; P.inpw = ctx->d_inpw64.p; P.inpb = ctx->d_inpb64.p;
    HC(hipMemsetAsync(ctx->d_F.p, 0, 3 * (size_t)ctx->N * sizeof(double), ctx->stream));
    HC(hipMemsetAsync(ctx->d_E.p, 0, ctx->S * sizeof(double), ctx->stream));
    for (int s = 0; s < ctx->S; ++s) {
        const int n = ctx->natoms[s], a0 = ctx->atom_off[s];
        const double* R = ctx->d_cell.p + 9 * s;
        if (timed(ctx, "plain_symm", [&] {
                k_plain_symm<<<cdiv(n, 64), 64, 0, ctx->stream>>>(n, a0, ctx->d_pmodel.p, P, ctx->d_pos.p, ctx->d_type.p, R,
                                                                   ctx->d_nb_off.p, ctx->d_nb_j.p, ctx->d_nb_img.p,
                                                                   ctx->d_G64.p);
            }))
            return -1;
        if (timed(ctx, "plain_force", [&] {
                k_plain_force<<<cdiv(n, 64), 64, 0, ctx->stream>>>(n, a0, ctx->d_pmodel.p, P, ctx->d_pos.p, ctx->d_type.p, R,
                                                                    ctx->d_nb_off.p, ctx->d_nb_j.p, ctx->d_nb_img.p,
                                                                    ctx->d_G64.p, ctx->d_F.p, ctx->d_E.p + s);
            }))
            return -1;
    }
    if (E_struc) HC(hipMemcpyAsync(E_struc, ctx->d_E.p, ctx->S * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    if (F) HC(hipMemcpyAsync(F, ctx->d_F.p, 3 * (size_t)ctx->N * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HC(hipStreamSynchronize(ctx->stream));
    if (ctx->nb_pending && nb_verify(ctx)) return -1;
    return flush_prof(ctx);
}

int ace_profile_enable(ace_ctx* ctx, int on) {
    if (!ctx) return -1;
    if (flush_prof(ctx)) return -1;
    ctx->prof = on != 0;
    if (on) ctx->acc.clear();
    return 0;
}

int ace_profile_read(const ace_ctx* cctx, const char* kernel, double* total_ms, long long* launches) {
    ace_ctx* ctx = const_cast<ace_ctx*>(cctx);
    if (!ctx) return -1;
    if (flush_prof(ctx)) return -1;
    auto it = ctx->acc.find(kernel ? kernel : "");
    if (it == ctx->acc.end()) {
        if (total_ms) *total_ms = 0;
        if (launches) *launches = 0;
        return 0;
    }
    if (total_ms) *total_ms = it->second.first;
    if (launches) *launches = it->second.second;
    return 0;
}

int ace_sync(ace_ctx* ctx) {
    if (!ctx) return -1;
    HC(hipStreamSynchronize(ctx->stream));
    if (ctx->nb_valid && nb_verify(ctx)) return -1;
    return flush_prof(ctx);
}

}  // extern "C"
