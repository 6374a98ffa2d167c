// basis_inst.hip -- one instantiation set of the basis kernels (basis.hip): real type ACE_INST_T
// (float or double, named ACE_INST_TN = f or d), ACE_INST_NG eta groups and ACE_INST_NZ power levels,
// modes 0-3, four and two waves per atom; with seven levels also the generic angular path.  Built once per
// combination by the Makefile.
#include <hip/hip_runtime.h>

#include "ace_common.hpp"
#include "basis.hip"
#include "basis_table.hpp"

#define ACE_CAT2(a, b, c, d) a##_##b##_##c##_##d
#define ACE_CAT(a, b, c, d) ACE_CAT2(a, b, c, d)

namespace {
template <class T, int NG, int NZ, int GEN>
void fill(BasisFn* f) {
    f[0] = k_basis<T, NG, NZ, 0, GEN>;
    f[1] = k_basis<T, NG, NZ, 1, GEN>;
    f[2] = k_basis<T, NG, NZ, 2, GEN>;
    f[3] = k_basis<T, NG, NZ, 3, GEN>;
    f[4] = k_basis<T, NG, NZ, 0, GEN, 2>;
    f[5] = k_basis<T, NG, NZ, 1, GEN, 2>;
    f[6] = k_basis<T, NG, NZ, 2, GEN, 2>;
    f[7] = k_basis<T, NG, NZ, 3, GEN, 2>;
}
}  // namespace

void ACE_CAT(basis_fill, ACE_INST_TN, ACE_INST_NG, ACE_INST_NZ)(BasisFn* f, bool gen) {
#if ACE_INST_NZ == 7
    if (gen) {
        fill<ACE_INST_T, ACE_INST_NG, ACE_INST_NZ, 1>(f);
        return;
    }
#endif
    (void)gen;
    fill<ACE_INST_T, ACE_INST_NG, ACE_INST_NZ, 0>(f);
}
