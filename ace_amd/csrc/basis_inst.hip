// basis_inst.hip -- one instantiation set of the basis kernels (basis.hip): real type ACE_INST_T
// (float or double, named ACE_INST_TN = f or d), ACE_INST_NG eta groups, ACE_INST_NZ power levels and the
// level split ACE_INST_SPL of two eta groups (0: none), modes 0-3, four and two waves per atom; with seven
// levels also the generic angular path.  Built once per combination by the Makefile.
#include <hip/hip_runtime.h>

#include "ace_common.hpp"
#include "basis.hip"
#include "basis_table.hpp"

#ifndef ACE_INST_SPL
#define ACE_INST_SPL 0
#endif

#define ACE_CAT2(a, b, c, d, e) a##_##b##_##c##_##d##_##e
#define ACE_CAT(a, b, c, d, e) ACE_CAT2(a, b, c, d, e)

namespace {
template <class T, int NG, int NZ, int GEN, int SPL>
void fill(BasisFn* f) {
    f[0] = k_basis<T, NG, NZ, 0, GEN, 4, SPL>;
    f[1] = k_basis<T, NG, NZ, 1, GEN, 4, SPL>;
    f[2] = k_basis<T, NG, NZ, 2, GEN, 4, SPL>;
    f[3] = k_basis<T, NG, NZ, 3, GEN, 4, SPL>;
    f[4] = k_basis<T, NG, NZ, 0, GEN, 2, SPL>;
    f[5] = k_basis<T, NG, NZ, 1, GEN, 2, SPL>;
    f[6] = k_basis<T, NG, NZ, 2, GEN, 2, SPL>;
    f[7] = k_basis<T, NG, NZ, 3, GEN, 2, SPL>;
}
}  // namespace

void ACE_CAT(basis_fill, ACE_INST_TN, ACE_INST_NG, ACE_INST_NZ, ACE_INST_SPL)(BasisFn* f, bool gen) {
#if ACE_INST_NZ == 7
    if (gen) {
        fill<ACE_INST_T, ACE_INST_NG, ACE_INST_NZ, 1, ACE_INST_SPL>(f);
        return;
    }
#endif
    (void)gen;
    fill<ACE_INST_T, ACE_INST_NG, ACE_INST_NZ, 0, ACE_INST_SPL>(f);
}
