/*
 * ace_oracle.h -- FP64 CPU restatement of the reference's Behler-Parrinello hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker the HIP path is compared against; only the tests,
 * the smoke check and the bench's CPU-baseline leg may load it.  Nothing in the product path links
 * or calls it.
 *
 * Each function cites the reference file:line it restates (paths relative to the ace source tree).
 * All arithmetic is IEEE FP64 with contraction disabled (-ffp-contract=off), in the reference's
 * loop order.  Eigen's GEMV/reduction blocking order is not reproduced (ulp level, unpinned).
 */
#ifndef ACE_ORACLE_H
#define ACE_ORACLE_H
#include "../include/ace_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* scalar math, src/math/special.cpp */
double or_coscut(double x);                 /* special.cpp:21-27  */
double or_sincut(double x);                 /* special.cpp:35-41  */
double or_fmexp(double x);                  /* special.cpp:249-284 */
double or_powint(double x, int n);          /* special.cpp:139-148 */

/* cutoff, src/nnp/cutoff.cpp:62-117 */
double or_cutf(int name, double rc, double r);
double or_cutg(int name, double rc, double r);

/* single basis functions, src/nnp/basis_radial.cpp:175-257, src/nnp/basis_angular.cpp:212-364 */
double or_radial_f(int form, int cut, double rc, double eta, double rs, double r);
double or_radial_d(int form, int cut, double rc, double eta, double rs, double r);
double or_angular_f(int form, int cut, double rc, double eta, double zeta, int lambda,
                    double cosv, double rij, double rik);
void   or_angular_d(int form, int cut, double rc, double eta, double zeta, int lambda,
                    double cosv, double rij, double rik, double* fphi, double* feta /*2*/);

/* model layout */
int or_input_width(const ace_model_desc* m, int I);        /* NNH::nInput */
int or_param_count(const ace_model_desc* m, int I);        /* ANN::size   */

/* NeighborList::build, src/struc/neighbor.cpp:46-87.  Returns the number of pairs; writes at most
 * cap pairs.  offsets has n+1 entries, img 3 ints per pair. */
long long or_neighbors(const double* R, int n, const double* pos, double rc, long long cap,
                       long long* offsets, int* jidx, int* img);

/* NNP::compute_symm, src/nnp/nnp.cpp:350-392.  G row per atom, stride ldg, zero-padded. */
void or_symm(const ace_model_desc* m, const double* R, int n, const double* pos, const int* type,
             const long long* offsets, const int* jidx, const int* img, int ldg, double* G);

/* ANN::fpbp + DODZ::grad for one atom of species I: returns out[0], fills dEdG (may be NULL).
 * src/ml/nn.cpp:1402-1420, 1836-1862 */
double or_network(const ace_model_desc* m, int I, const double* G, double* dEdG);

/* NNP::compute_energy, src/nnp/nnp.cpp:329-338 (+ NNH::energy nnh.cpp:207-210): sum(out + E0). */
double or_energy(const ace_model_desc* m, int n, const int* type, int ldg, const double* G);

/* NNP::compute_force, src/nnp/nnp.cpp:412-476: fills F (n x 3), returns sum(out) (no E0).
 * virial (9, row-major W_ab = -sum_pairs (dE/dd)_a d_b) may be NULL. */
double or_force(const ace_model_desc* m, const double* R, int n, const double* pos, const int* type,
                const long long* offsets, const int* jidx, const int* img, int ldg,
                const double* G, double* F, double* virial);

/*
 * Training loss and gradient of one structure (accumulated into grad, all species concatenated).
 * Energy term: NNPTEFR::error_cost_o1 (src/nnp/nnptefr.cpp:1336-1382) with E = sum(out+E0),
 * dU = (E - E_ref) * norm(n), contribution w_energy * loss(dU), gradient w_energy * loss'(dU) *
 * norm * sum_i dE_i/dW (NN::Cost::grad, src/ml/nn.cpp:1747-1808).
 * Force term (no reference counterpart; network factor is NN::D2ODZDI, src/ml/nn.cpp:2047-2140):
 * w_force * sum_a 0.5 |F_a - F_ref_a|^2, gradient by forward-over-reverse through the network with
 * dG/dr recomputed (per-function derivatives of basis_radial.cpp:210-257, basis_angular.cpp:255-364).
 * out_loss[0] = energy term, out_loss[1] = force term, out_loss[2] = (dE/n)^2.
 */
void or_train_struct(const ace_model_desc* m, const ace_loss_desc* L, const double* R, int n,
                     const double* pos, const int* type, const long long* offsets, const int* jidx,
                     const int* img, int ldg, const double* G, double E_ref, const double* F_ref,
                     double* grad, double* out_loss);

/* ADAM::step, src/opt/algo.cpp:280-296 (beta1 0.9, beta2 0.999, algo.cpp:260-261).
 * state: m (dim), v (dim), b[0] = beta1^t, b[1] = beta2^t (start at beta1, beta2). */
void or_adam_step(int dim, double* p, const double* g, double gamma, double eps, double* m,
                  double* v, double* b);

#ifdef __cplusplus
}
#endif
#endif
