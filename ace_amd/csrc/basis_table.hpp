// basis_table.hpp -- the launch table of the basis kernels.  Each (real type, eta groups, power levels,
// level split) combination is instantiated in its own translation unit (basis_inst.hip, built once per
// combination), so the kernels compile in parallel; the energy/force code picks its table through
// basis_fill().
#pragma once

typedef void (*BasisFn)(int, const DevModel*, DevBasisParams, BasisIO, int, int, int);

// f[0..3]: modes 0-3 with four waves per atom; f[4..7]: modes 0-3 with two.  gen: the model has angular
// blocks outside the canonical (lambda = +-1, zeta = 2^k) form, compiled only with seven power levels.
#define ACE_BASIS_FILL(TN, NG, NZ, S) void basis_fill_##TN##_##NG##_##NZ##_##S(BasisFn* f, bool gen);
#define ACE_BASIS_FILL_NZ(TN, NG) ACE_BASIS_FILL(TN, NG, 3, 0) ACE_BASIS_FILL(TN, NG, 4, 0) \
    ACE_BASIS_FILL(TN, NG, 5, 0) ACE_BASIS_FILL(TN, NG, 7, 0)
ACE_BASIS_FILL_NZ(f, 1)
ACE_BASIS_FILL_NZ(f, 2)
ACE_BASIS_FILL_NZ(f, 4)
ACE_BASIS_FILL_NZ(d, 1)
ACE_BASIS_FILL_NZ(d, 2)
ACE_BASIS_FILL_NZ(d, 4)
// FP32 kernels of two eta groups with a level split (group 0 below level S, group 1 from S on)
ACE_BASIS_FILL(f, 2, 3, 1) ACE_BASIS_FILL(f, 2, 3, 2)
ACE_BASIS_FILL(f, 2, 4, 1) ACE_BASIS_FILL(f, 2, 4, 2) ACE_BASIS_FILL(f, 2, 4, 3)
ACE_BASIS_FILL(f, 2, 5, 1) ACE_BASIS_FILL(f, 2, 5, 2) ACE_BASIS_FILL(f, 2, 5, 3) ACE_BASIS_FILL(f, 2, 5, 4)
ACE_BASIS_FILL(f, 2, 7, 1) ACE_BASIS_FILL(f, 2, 7, 2) ACE_BASIS_FILL(f, 2, 7, 3) ACE_BASIS_FILL(f, 2, 7, 4)
ACE_BASIS_FILL(f, 2, 7, 5) ACE_BASIS_FILL(f, 2, 7, 6)
#undef ACE_BASIS_FILL_NZ
#undef ACE_BASIS_FILL

// power levels compiled for a model whose canonical blocks use nz levels (zeta up to 2^(nz-1))
inline int basis_levels(int nz, bool gen) {
    if (gen || nz > 5) return 7;
    return nz <= 3 ? 3 : nz;
}

// the table for real type f64, ng eta groups (1, 2 or 4), nz levels and the level split spl of two eta
// groups (FP32 only; 0: none)
inline void basis_fill(bool f64, int ng, int nz, bool gen, int spl, BasisFn* f) {
    const int z = basis_levels(nz, gen);
    if (!f64 && ng == 2 && spl > 0 && spl < z) {
        switch (z * 10 + spl) {
            case 31: basis_fill_f_2_3_1(f, gen); return;
            case 32: basis_fill_f_2_3_2(f, gen); return;
            case 41: basis_fill_f_2_4_1(f, gen); return;
            case 42: basis_fill_f_2_4_2(f, gen); return;
            case 43: basis_fill_f_2_4_3(f, gen); return;
            case 51: basis_fill_f_2_5_1(f, gen); return;
            case 52: basis_fill_f_2_5_2(f, gen); return;
            case 53: basis_fill_f_2_5_3(f, gen); return;
            case 54: basis_fill_f_2_5_4(f, gen); return;
            case 71: basis_fill_f_2_7_1(f, gen); return;
            case 72: basis_fill_f_2_7_2(f, gen); return;
            case 73: basis_fill_f_2_7_3(f, gen); return;
            case 74: basis_fill_f_2_7_4(f, gen); return;
            case 75: basis_fill_f_2_7_5(f, gen); return;
            default: basis_fill_f_2_7_6(f, gen); return;
        }
    }
#define ACE_PICK(TN)                                                        \
    switch (ng * 10 + z) {                                                  \
        case 13: basis_fill_##TN##_1_3_0(f, gen); break;                    \
        case 14: basis_fill_##TN##_1_4_0(f, gen); break;                    \
        case 15: basis_fill_##TN##_1_5_0(f, gen); break;                    \
        case 17: basis_fill_##TN##_1_7_0(f, gen); break;                    \
        case 23: basis_fill_##TN##_2_3_0(f, gen); break;                    \
        case 24: basis_fill_##TN##_2_4_0(f, gen); break;                    \
        case 25: basis_fill_##TN##_2_5_0(f, gen); break;                    \
        case 27: basis_fill_##TN##_2_7_0(f, gen); break;                    \
        case 43: basis_fill_##TN##_4_3_0(f, gen); break;                    \
        case 44: basis_fill_##TN##_4_4_0(f, gen); break;                    \
        case 45: basis_fill_##TN##_4_5_0(f, gen); break;                    \
        default: basis_fill_##TN##_4_7_0(f, gen); break;                    \
    }
    if (f64) {
        ACE_PICK(d)
    } else {
        ACE_PICK(f)
    }
#undef ACE_PICK
}
