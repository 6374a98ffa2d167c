/*
 * ace_hip.h -- C-ABI of the MI355X Behler-Parrinello energy/force/training path.
 *
 * Every entry point is plain C: borrowed host pointers, sizes, an int status
 * (0 = success, < 0 = failure; the message is in ace_last_error()).  No C++
 * exception, torch type or HIP type crosses this boundary.
 *
 * Reference interfaces replaced (paths relative to the ace source tree):
 *   ace_model_desc / ace_create ... NNP + NNH + BasisR/BasisA + NN::ANN set-up
 *                                   (src/nnp/nnp.hpp:64-121, src/nnp/nnh.hpp:46-155,
 *                                    src/nnp/basis_radial.hpp:23-89,
 *                                    src/nnp/basis_angular.hpp:23-103, src/ml/nn.hpp:243-369)
 *   ace_set_params / ace_get_params NN::ANN operator<< / operator>> (src/ml/nn.cpp:881-917)
 *   ace_upload ..................... Structure positions/types/cell (src/struc/structure.hpp:147-164)
 *   ace_build_neighbors ............ NeighborList::build (src/struc/neighbor.cpp:46-87)
 *   ace_export_neighbors ........... NeighborList::index/img accessors (src/struc/neighbor.hpp:34-40)
 *   ace_symm ....................... NNP::compute_symm (src/nnp/nnp.cpp:350-392)
 *   ace_energy_forces .............. NNP::compute / compute_force / compute_energy
 *                                   (src/nnp/nnp.cpp:329-338, 412-494)
 *   ace_compute .................... NNP::compute(Structure&) (src/nnp/nnp.cpp:489-494): build + energy
 *   ace_energy_forces_noalloc ...... (no counterpart) ace_energy_forces into device memory, allocating nothing
 *   ace_train_step ................. NNPTEFR::error_cost_o1 (src/nnp/nnptefr.cpp:1292-1425)
 *                                    extended with the force-matching term d(F)/dW
 *                                    (network factor NN::D2ODZDI, src/ml/nn.cpp:2047-2140)
 *   ace_train_optimizer ............ (no counterpart) the ADAM update of the training loop
 *                                    (nnptefr.cpp:1031-1066, opt::ADAM algo.cpp:280-296) on the GPU
 *   ace_train_opt_state ............ (no counterpart) its state, for restart files
 *   ace_forward / ace_backward ..... autograd-style pair: the energy/force pass, then d(loss)/d(params) from
 *                                    the caller's dL/dE and dL/dF (the network factor NN::D2ODZDI::grad,
 *                                    src/ml/nn.cpp:2047-2140, seeded as NNPTEFR::error_cost_o1 seeds it,
 *                                    nnptefr.cpp:1353-1378, with the upstream vectors in place of the loss's)
 *   ace_md_init / ace_md_run / ace_md_state  velocity-Verlet steps (src/sim/integrator.cpp:548-571), HIP graphs
 *   ace_autotune ................... (no counterpart) measured choice of the energy/force path and of
 *                                    its tile shapes
 *   ace_tile_shape ................. (no counterpart) the shapes in use
 *   ace_set_tile_shape ............. (no counterpart) fixed tile shapes (chunk weight, network waves)
 *   ace_set_level_split ............ (no counterpart) level-split angular kernels on or off
 *   ace_set_force_gather ........... (no counterpart) forces gathered from per-pair gradients, or
 *                                    fixed-point atomics per pair
 *   ace_set_path ................... (no counterpart) fixed choice of the energy/force path
 *   ace_plain_energy_forces ........ same as ace_energy_forces, unbatched FP64 one-thread-per-atom
 *                                   GPU port (the baseline the batched path is measured against)
 */
#ifndef ACE_HIP_H
#define ACE_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

#define ACE_MAX_TYPES 8
#define ACE_MAX_LAYERS 6   /* weight layers per network, output layer included */

/* enum ids follow the reference's enum orders */
enum { ACE_CUT_NONE = 0, ACE_CUT_STEP = 1, ACE_CUT_COS = 2, ACE_CUT_TANH = 3, ACE_CUT_POLY3 = 4 };
enum { ACE_RAD_NONE = 0, ACE_RAD_GAUSSIAN = 1, ACE_RAD_SECH = 2, ACE_RAD_LOGISTIC = 3,
       ACE_RAD_TANH = 4, ACE_RAD_LOGCOSH = 5, ACE_RAD_LOGCOSH2 = 6 };
enum { ACE_ANG_NONE = 0, ACE_ANG_GAUSS = 1, ACE_ANG_GAUSS2 = 2, ACE_ANG_SECH = 3,
       ACE_ANG_STUDENT3 = 4, ACE_ANG_STUDENT4 = 5, ACE_ANG_STUDENT5 = 6 };
enum { ACE_ACT_LINEAR = 1, ACE_ACT_SIGMOID = 2, ACE_ACT_TANH = 3, ACE_ACT_SWISH = 10,
       ACE_ACT_SOFTPLUS = 15 };
/* ACE_PREC_FP32: symmetry functions, networks and adjoints in FP32 (FP64 reductions).
 * ACE_PREC_FP64_BASIS: symmetry functions and their adjoint/tangent in FP64, networks in FP32.
 * ACE_PREC_BF16L1 (ace_energy_forces / ace_compute only): as FP32, with the first network layer's forward
 *   product and its dE/dG product on BF16 matrix cores with FP32 sums; each operand is split into BF16
 *   high and low parts and three products are summed (hi*hi + hi*lo + lo*hi). */
enum { ACE_PREC_FP32 = 0, ACE_PREC_BF16L1 = 1, ACE_PREC_FP64_BASIS = 2 };
enum { ACE_LOSS_MSE = 0, ACE_LOSS_MAE = 1, ACE_LOSS_HUBER = 2, ACE_LOSS_ASINH = 3 };
enum { ACE_NORM_IDENTITY = 0, ACE_NORM_LINEAR = 1, ACE_NORM_SQRT = 2, ACE_NORM_CBRT = 3,
       ACE_NORM_LOG = 4 };

/*
 * Model description: the content of an NNP (one NNH per species).
 *   radial block   (I,J)   -> index I*ntypes + J             (NNH::basisR(J))
 *   angular block  (I,J,K) -> index I*nab + b, nab = ntypes*(ntypes+1)/2,
 *                             b = max(J,K)*(max(J,K)+1)/2 + min(J,K)
 *                             (TensorTri<2> order, src/nnp/tensortri.hpp:99-111)
 *   input vector of species I = [radial blocks J=0..][angular blocks b=0..]
 *                             (NNH::init_input, src/nnp/nnh.cpp:183-201)
 *   network of species I: nn_nlayer[I] weight layers, widths
 *     nn_width[I*(ACE_MAX_LAYERS+1) + l], l = 0..nlayer (width 0 = inputs, last = 1);
 *     hidden activation nn_act[I] with sharpness nn_c[I], output layer linear
 *     (src/ml/nn.cpp:1349-1352); ins = inpw*(G+inpb), out = outb + outw*a_last.
 *   params of species I start at par_off[I]: all biases b_0..b_{L-1}, then all
 *     weights W_0..W_{L-1} column-major (src/ml/nn.cpp:881-896).
 */
typedef struct ace_model_desc {
    int ntypes;
    double rc;                 /* neighbour-list cutoff (NNP::rc_) */
    /* radial */
    const int*    r_form;
    const int*    r_cut;
    const double* r_rc;
    const int*    r_n;
    const int*    r_off;       /* offset of the block in r_rs / r_eta */
    const double* r_rs;
    const double* r_eta;
    /* angular */
    const int*    a_form;
    const int*    a_cut;
    const double* a_rc;
    const int*    a_n;
    const int*    a_off;       /* offset of the block in a_eta / a_zeta / a_lambda */
    const double* a_eta;
    const double* a_zeta;
    const int*    a_lambda;
    /* networks */
    const int*    nn_nlayer;
    const int*    nn_width;
    const int*    nn_act;
    const double* nn_c;
    const int*    in_off;      /* offset of species I in inpw / inpb */
    const double* inpw;
    const double* inpb;
    const double* outw;        /* per species */
    const double* outb;        /* per species */
    const double* E0;          /* per species (NNH::type().energy()) */
    const int*    par_off;     /* offset of species I in params */
    const double* params;
} ace_model_desc;

/* Training-loss description (energy term as NNPTEFR::error_cost_o1; force term new). */
typedef struct ace_loss_desc {
    int loss;          /* ACE_LOSS_* applied to the normalised energy error dU */
    int norm;          /* ACE_NORM_* : dU = (E - E_ref) * norm(nAtoms) */
    double delta;      /* HUBER / ASINH scale */
    double w_energy;   /* multiplies the energy term; the reference divides by the batch size */
    double w_force;    /* multiplies sum over atoms,xyz of 0.5*(F - F_ref)^2 */
} ace_loss_desc;

typedef struct ace_ctx ace_ctx;

/* Context: device buffers, one HIP stream, model copy.  One context per host thread. */
int  ace_create(ace_ctx** ctx, const ace_model_desc* model);
void ace_destroy(ace_ctx* ctx);
const char* ace_last_error(const ace_ctx* ctx);
int  ace_input_width(const ace_ctx* ctx, int type);          /* NNH::nInput */
int  ace_param_count(const ace_ctx* ctx, int type);          /* NN::ANN::size */

/* NN::ANN operator<< / operator>> for species `type` (parameter order of nn.cpp:881-896). */
int  ace_set_params(ace_ctx* ctx, int type, const double* params, int n);
int  ace_get_params(const ace_ctx* ctx, int type, double* params, int n);
/* input preconditioning (ANN::inpw / inpb), as NNPTEFR::train sets it (nnptefr.cpp:786-880) */
int  ace_set_input_scaling(ace_ctx* ctx, int type, const double* inpw, const double* inpb, int n);

/*
 * Structures: n_struc structures, atom counts natoms[s], positions pos (sum(natoms) x 3, Angstrom,
 * structure after structure), species type (sum(natoms)), cells cell (n_struc x 9, column-major,
 * columns = lattice vectors, as Structure::R()).  Copied to HBM; kept until the next upload.
 */
int  ace_upload(ace_ctx* ctx, int n_struc, const int* natoms, const double* pos, const int* type,
                const double* cell);
/* Update positions only (same structures, same atom order); for MD and finite differences. */
int  ace_update_positions(ace_ctx* ctx, const double* pos);

/*
 * Neighbour list for every uploaded structure: all (j, n) with |r_i - r_j - R n|^2 < (rc+skin)^2,
 * j = i allowed only for n != 0, n within the reference's image range ceil(3 rc / |R e_d|)
 * (NeighborList::build, neighbor.cpp:46-87).  Decided in FP64 with the reference formula.
 */
int  ace_build_neighbors(ace_ctx* ctx, double skin);
int  ace_neighbor_count(const ace_ctx* ctx, long long* npairs);
/* offsets: sum(natoms)+1 (global atom index, per-atom list start); j: local atom index within the
 * structure; img: 3 ints per pair.  Lists are in the reference order: j ascending, then n
 * lexicographic. */
int  ace_export_neighbors(ace_ctx* ctx, long long* offsets, int* j, int* img);

/* Precision of ace_symm and ace_train_step (ACE_PREC_*; default ACE_PREC_FP32).  No reference
 * counterpart: the reference computes everything in FP64. */
int  ace_set_precision(ace_ctx* ctx, int precision);

/* Symmetry functions G (on the GPU, stored FP32); G exported row per atom with stride ldg (>= max
 * width), zero-padded, as FP64.  G may be NULL (computed and kept on the device). */
int  ace_symm(ace_ctx* ctx, double* G, int ldg);

/*
 * Energy and forces of every uploaded structure in one batched pass (neighbour list must be built):
 *   E_struc[s] = sum_i (NN_i + E0_i)            (NNP::compute_energy convention, FP64 reduction)
 *   F          = -dE/dr, sum(natoms) x 3         (NNP::compute_force)
 *   virial     = n_struc x 9, W_ab = -sum_pairs d_a dE/dd_b (FP64 reduction); may be NULL
 * precision: ACE_PREC_FP32, ACE_PREC_FP64_BASIS or ACE_PREC_BF16L1.  Output pointers may be NULL: results
 * stay in HBM.
 */
int  ace_energy_forces(ace_ctx* ctx, int precision, double* E_struc, double* F, double* virial);

/* NNP::compute(Structure&) (src/nnp/nnp.cpp:489-494): neighbour list with list radius rc + skin, then
 * ace_energy_forces, in one call. */
int  ace_compute(ace_ctx* ctx, int precision, double skin, double* E_struc, double* F, double* virial);

/* No counterpart in ace: for callers that cannot allocate device memory (code under HIP graph capture,
 * serving loops, harnesses that forbid hipMalloc).  Same quantities as ace_energy_forces (same kernels, same
 * bits) for the uploaded batch and its built neighbour list, with every pointer a DEVICE pointer: pos (N x 3,
 * may be NULL: the uploaded positions) is copied in first, E_struc (n_struc), F (N x 3) and virial
 * (n_struc x 9, may be NULL) are written.  The call allocates nothing and does not synchronise: its
 * accumulators, per-atom energies and network rows live in one static 64 MiB device array per GPU (about
 * 56 bytes per atom with the fused kernel, plus 8 * ldg bytes per atom with the split path; a larger batch
 * is refused with a message).  Calls that overlap on one GPU share that array and race.  Parameters or
 * input scaling changed since the last ace_energy_forces / ace_symm call are refused (those calls upload
 * them).  An atom that outgrew the LDS slots is reported by the next synchronising call. */
int  ace_energy_forces_noalloc(ace_ctx* ctx, int precision, const double* pos, double* E_struc, double* F,
                               double* virial);

/* Velocity-Verlet molecular dynamics of the uploaded batch, metal units (A, ps, amu, eV).  No counterpart in
 * ace's NNP API: ace has no NNP molecular dynamics; the step follows its Verlet template
 * (src/sim/integrator.cpp:548-571: half kick, drift, forces, half kick), with a = F / (m * 1.0364269e-4).
 * ace_md_init: per-species masses (amu) and velocities (N x 3, A/ps; NULL = zero).
 * ace_md_run: nsteps steps of dt (ps).  The neighbour list (radius rc + skin) is rebuilt before the first
 *   step and every rebuild_every steps, outside any graph; with use_graph != 0 the steps in between are
 *   replayed from a captured HIP graph of two steps (bit-identical to use_graph = 0).  energy: NULL or
 *   nsteps x 2 (potential energy of the batch with E0, kinetic energy; eV) after each step.  Positions and
 *   velocities stay on the device.
 * ace_md_state: current positions and velocities (either pointer may be NULL). */
int  ace_md_init(ace_ctx* ctx, const double* mass, const double* vel);
int  ace_md_run(ace_ctx* ctx, int precision, int nsteps, double dt, int rebuild_every, double skin, int use_graph,
                double* energy);
int  ace_md_state(ace_ctx* ctx, double* pos, double* vel);

/* Energy/force paths: the fused per-atom kernel (symmetry functions, network, adjoint) with four or two
 * waves per atom, or symmetry functions -> network -> adjoint with the network weights staged in LDS or
 * read from memory, its per-atom kernels with four or two waves per atom. */
enum { ACE_PATH_FUSED = 0, ACE_PATH_SPLIT_STAGED = 1, ACE_PATH_SPLIT = 2, ACE_PATH_FUSED_2W = 3,
       ACE_PATH_SPLIT_STAGED_2W = 4, ACE_PATH_SPLIT_2W = 5, ACE_NPATHS = 6 };
/* Autotuner (no counterpart in ace): times the paths on the uploaded batch (neighbour list built) and
 * keeps the fastest for this precision, then, on that path, the tile shape of the angular triplet loops:
 * the chunk of partner slots a lane takes per fixed slot, chosen per block to minimise
 * rounds x (chunk + w/8) with w in {1, 3, 8} (w: the cost of a chunk's set-up and flush in eighths of one
 * pair; a larger w favours long chunks, a smaller one fuller rounds), and on a split path the waves that
 * share a 16-atom network tile (1, 2 or 4).  ms: NULL or ACE_NPATHS + 1 entries,
 * the milliseconds of each path (-1 where not feasible) and the index chosen.  The choice holds while the
 * batch size and neighbour-row capacity stay as measured; otherwise size rules decide. */
int  ace_autotune(ace_ctx* ctx, int precision, double* ms);
/* The shapes the energy/force pass of this precision uses now: shape[0] path (ACE_PATH_*, -1: size rule),
 * shape[1] waves per atom of its basis kernels, shape[2] network weights staged in LDS (1) or not (0),
 * shape[3] the angular chunk weight w, shape[4] the milliseconds of the chosen chunk weight's timing (or -1),
 * shape[5..7] the milliseconds measured for w = 1, 3, 8 (-1 where not measured), shape[8] the level split s
 * of the canonical angular sums (0: none; with two eta groups whose power levels do not overlap, group 0
 * below level s and group 1 from s on, the FP32 kernels form only those (level, group) products),
 * shape[9] the waves per 16-atom network tile of the split path (1, 2 or 4; 0 on a fused path),
 * shape[10..12] the milliseconds measured with 1, 2, 4 waves per tile (-1 where not measured).  shape holds
 * 13 entries. */
int  ace_tile_shape(const ace_ctx* ctx, int precision, double* shape);
/* (no counterpart) Fixes tile shapes of a precision over the autotuner's choice and the defaults: the
 * chunk weight w of the angular triplet loops (1, 3 or 8) and the waves per 16-atom network tile (1, 2 or
 * 4); 0 returns either to the tuned or default value. */
int  ace_set_tile_shape(ace_ctx* ctx, int precision, int chunk_w, int net_waves);
/* (no counterpart) Enables (default) or disables the level-split kernels; both give the oracle's results
 * within the FP32 bounds, in a different rounding order. */
int  ace_set_level_split(ace_ctx* ctx, int enable);
/* (no counterpart) Forces of the split energy/force passes: 1 (default) stores each pair's gradient
 * dE_i/dd at the key-row entry of its reverse pair (3 doubles per entry and 3 per atom, allocated by
 * ace_build_neighbors) and sums each row in a fixed order; 0 adds them with 64-bit fixed-point atomics as
 * the fused kernel and training do.  Both scatter as NNP::compute_force (nnp.cpp:442-471) and agree to
 * ~1e-9 eV/A; each is deterministic.  Takes effect from the next ace_build_neighbors when turned on.  The
 * FP64-basis mode always uses the atomics. */
int  ace_set_force_gather(ace_ctx* ctx, int enable);
/* Fixes the energy/force path of a precision (ACE_PATH_*) for every batch, the wave count of the
 * symmetry-function, adjoint and tangent kernels of training included (a fused path whose workgroup does
 * not fit in LDS runs split); -1 returns to the size rules.  The BF16 first-layer mode has the split
 * paths only. */
int  ace_set_path(ace_ctx* ctx, int precision, int path);

/* Same quantities from the unbatched FP64 port: one structure per launch, one thread per atom. */
int  ace_plain_energy_forces(ace_ctx* ctx, double* E_struc, double* F);

/*
 * One training step over all uploaded structures (the batch): loss and d(loss)/d(params) for every
 * species, params concatenated over species (NNPTEFR packs gElement_ the same way,
 * nnptefr.cpp:983-987).  E_ref per structure (total energy incl. E0), F_ref sum(natoms) x 3
 * (may be NULL when w_force == 0).  loss[0] = total, loss[1] = energy part, loss[2] = force part,
 * loss[3] = sum_s (dE/nAtoms)^2.  grad may be NULL.
 */
int  ace_set_targets(ace_ctx* ctx, const double* E_ref, const double* F_ref);
/* on != 0: ace_train_step replays its device launches (networks, forces, loss, tangent, gradient) from a
 * captured HIP graph (no counterpart in ace; same results as the plain launches). */
int  ace_train_graph(ace_ctx* ctx, int on);
int  ace_train_step(ace_ctx* ctx, const ace_loss_desc* loss_desc, double* loss, double* grad);

/*
 * Autograd-style boundary over the same kernels as ace_train_step.  ace_forward: energies E_struc[S] (with
 * E0), forces F[3N] and virials[9S] at the current parameters and positions, in the precision of
 * ace_set_precision (any output may be NULL); it keeps the network-input rows for ace_backward.
 * ace_backward: with upstream gradients dL/dE_struc[S] and dL/dF[3N] (either may be NULL, meaning zero) of a
 * caller's loss L(E, F), writes dL/dparams (the parameter count of every species, concatenated in
 * ace_get_params order) and dL/dpos[3N] = -dL/dE_s F_a (the energy term; with dL_dF given, dL/dpos would need
 * the Hessian-vector product, which is not computed, and the call fails).  Either output may be NULL.  Valid
 * after ace_forward until positions, parameters, input scaling or precision change or another device pass
 * (ace_energy_forces, ace_train_step, ...) runs; it may be called several times per forward.  For the seeds of
 * ace_train_step's loss (dL/dE_s = w_energy nf(N_s) dU_s, dL/dF = w_force (F - F_ref)) dL/dparams equals
 * ace_train_step's gradient.
 */
int  ace_forward(ace_ctx* ctx, double* E_struc, double* F, double* virial);
int  ace_backward(ace_ctx* ctx, const double* dL_dE, const double* dL_dF, double* dL_dparams, double* dL_dpos);

/* Device optimiser (no counterpart in ace, whose training loop updates the parameters on the host,
 * nnptefr.cpp:1031-1066).  algo ACE_OPT_ADAM: every later ace_train_step also applies one ADAM step
 * (opt::ADAM, algo.cpp:280-296; beta1 0.9, beta2 0.999, constant gamma, eps) to FP64 master parameters on
 * the GPU and refreshes the networks' device layouts in the same stream, so a training loop moves no
 * parameters between host and device; ace_get_params reads the current parameters back, and parameters
 * set with ace_set_params become the master copy.  Starts from the current parameters with zero moments.
 * ACE_OPT_HOST: back to host updates (the caller's ace_set_params after each step). */
enum { ACE_OPT_HOST = 0, ACE_OPT_ADAM = 1 };
int  ace_train_optimizer(ace_ctx* ctx, int algo, double gamma, double eps);
/* The device optimiser's state: set == 0 reads, set != 0 writes p, m, v (the parameter count each,
 * species concatenated) and bt[2] = (beta1^t, beta2^t) of the next step. */
int  ace_train_opt_state(ace_ctx* ctx, int set, double* p, double* m, double* v, double* bt);

/* Timing of the last calls: per-kernel HIP-event durations accumulated while enabled. */
int  ace_profile_enable(ace_ctx* ctx, int on);
int  ace_profile_read(const ace_ctx* ctx, const char* kernel, double* total_ms, long long* launches);
int  ace_sync(ace_ctx* ctx);

#ifdef __cplusplus
}
#endif
#endif
