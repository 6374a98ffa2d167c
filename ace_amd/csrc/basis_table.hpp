// basis_table.hpp -- the launch table of the basis kernels.  Each (real type, eta groups, power levels)
// combination is instantiated in its own translation unit (basis_inst.hip, built once per combination), so
// the 240 kernels compile in parallel; the energy/force code picks its table through basis_fill().
#pragma once

typedef void (*BasisFn)(int, const DevModel*, DevBasisParams, BasisIO, int, int, int);

// f[0..3]: modes 0-3 with four waves per atom; f[4..7]: modes 0-3 with two.  gen: the model has angular
// blocks outside the canonical (lambda = +-1, zeta = 2^k) form, compiled only with seven power levels.
#define ACE_BASIS_FILL(TN, NG, NZ) void basis_fill_##TN##_##NG##_##NZ(BasisFn* f, bool gen);
#define ACE_BASIS_FILL_NZ(TN, NG) ACE_BASIS_FILL(TN, NG, 3) ACE_BASIS_FILL(TN, NG, 4) ACE_BASIS_FILL(TN, NG, 5) \
    ACE_BASIS_FILL(TN, NG, 7)
ACE_BASIS_FILL_NZ(f, 1)
ACE_BASIS_FILL_NZ(f, 2)
ACE_BASIS_FILL_NZ(f, 4)
ACE_BASIS_FILL_NZ(d, 1)
ACE_BASIS_FILL_NZ(d, 2)
ACE_BASIS_FILL_NZ(d, 4)
#undef ACE_BASIS_FILL_NZ
#undef ACE_BASIS_FILL

// power levels compiled for a model whose canonical blocks use nz levels (zeta up to 2^(nz-1))
inline int basis_levels(int nz, bool gen) {
    if (gen || nz > 5) return 7;
    return nz <= 3 ? 3 : nz;
}

// the table for real type f64, ng eta groups (1, 2 or 4) and nz levels
inline void basis_fill(bool f64, int ng, int nz, bool gen, BasisFn* f) {
    const int z = basis_levels(nz, gen);
#define ACE_PICK(TN)                                                        \
    switch (ng * 10 + z) {                                                  \
        case 13: basis_fill_##TN##_1_3(f, gen); break;                      \
        case 14: basis_fill_##TN##_1_4(f, gen); break;                      \
        case 15: basis_fill_##TN##_1_5(f, gen); break;                      \
        case 17: basis_fill_##TN##_1_7(f, gen); break;                      \
        case 23: basis_fill_##TN##_2_3(f, gen); break;                      \
        case 24: basis_fill_##TN##_2_4(f, gen); break;                      \
        case 25: basis_fill_##TN##_2_5(f, gen); break;                      \
        case 27: basis_fill_##TN##_2_7(f, gen); break;                      \
        case 43: basis_fill_##TN##_4_3(f, gen); break;                      \
        case 44: basis_fill_##TN##_4_4(f, gen); break;                      \
        case 45: basis_fill_##TN##_4_5(f, gen); break;                      \
        default: basis_fill_##TN##_4_7(f, gen); break;                      \
    }
    if (f64) {
        ACE_PICK(d)
    } else {
        ACE_PICK(f)
    }
#undef ACE_PICK
}
