// tdda_zy_plan.hpp -- lengths, registers and LDS of the fused z + y passes
// (tdda_zy.hpp), for the planner and the kernels.
#pragma once
#include "tdda_fft.hpp"

namespace tdda {

// z lengths of the fused passes (each a register DFT of sdft)
#define TDDA_ZY_ZLENGTHS(X) X(8) X(9) X(10) X(12) X(15) X(16) X(18) X(20) X(24) X(25) X(27) X(30) X(32)

// Longest z length of the fused passes per precision: complex128 holds
// twice the registers per point, and its inverse z transform of more than
// TDDA_ZY_ZMAX_C128 points would spill.
#ifndef TDDA_ZY_ZMAX_C128
#define TDDA_ZY_ZMAX_C128 20
#endif
template <typename T>
__host__ __device__ constexpr int zy_zmax()
{
    return sizeof(T) == 8 ? TDDA_ZY_ZMAX_C128 : 32;
}

template <typename T>
__host__ __device__ constexpr bool zy_zlen_ok(int L)
{
    return L <= zy_zmax<T>() && (L == 8 || L == 9 || L == 10 || L == 12 || L == 15 || L == 16 || L == 18 ||
                                 L == 20 || L == 24 || L == 25 || L == 27 || L == 30 || L == 32);
}

// y lengths of the fused passes: the register four-step lengths up to 128
// (factors up to 16, 256 threads)
#define TDDA_ZY_YLENGTHS(X)                                                                                 \
    X(32) X(36) X(40) X(45) X(48) X(50) X(54) X(60) X(64) X(72) X(75) X(80) X(81) X(90) X(96) X(100)     \
        X(108) X(120) X(128)

__host__ __device__ constexpr bool zy_ylen_ok(int L)
{
    return L == 32 || L == 36 || L == 40 || L == 45 || L == 48 || L == 50 || L == 54 || L == 60 || L == 64 ||
           L == 72 || L == 75 || L == 80 || L == 81 || L == 90 || L == 96 || L == 100 || L == 108 || L == 120 ||
           L == 128;
}

// Waves per SIMD the fused kernels are compiled for: the inverse z
// transform holds all Mz <= 32 points of a pencil (the forward one half of
// them, the rest is zero padding) plus the split's temporaries.
template <typename T, bool INV>
__host__ __device__ constexpr int zy_waves()
{
    return sizeof(T) == 8 ? 2 : (INV ? 3 : 4);
}

// Row stride of the LDS slab [y][kz][c]: 3 Mz plus one spare slot (odd, so
// the rows of one wave's column reads fall on different banks).
__host__ __device__ constexpr int zy_row_stride(int Mz) { return 3 * Mz + 1; }

// LDS slots of a fused tile: y twiddles, the slab, and one region shared by
// the staged inputs / z outputs (Ny Nz 3) and the y exchange (nf columns).
__host__ __device__ constexpr size_t zy_lds_slots(int LY, int Ny, int Nz, int Mz, int nf, int ps)
{
    const size_t a = (size_t)Ny * Nz * 3, b = (size_t)nf * ps;
    return (size_t)pass_tw_slots(true, LY, 0) + (size_t)Ny * zy_row_stride(Mz) + (a > b ? a : b);
}

}  // namespace tdda
