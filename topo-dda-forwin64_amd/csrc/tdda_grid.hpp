// tdda_grid.hpp -- default zero-padded FFT grid of an operator (host only).
//
// Any M >= 2N - 1 along an axis gives the same linear convolution.  The y and
// z lengths set the bytes of every pass (the intermediates hold My Mz
// points per x row), so they stay the smallest 2,3,5-smooth length, except
// where that length is known slow.  The x length sets only arithmetic: the
// x pass loads and stores the Nx rows that exist and transforms them in
// registers or LDS, so its cost is that of the transform, which varies with
// the factors of the length more than with the length itself (125 = 5^3
// costs 1.5x the arithmetic of 128 in complex64 and, without a register
// kernel, 2.4x in complex128).  The default x length is therefore chosen
// among the lengths within 30 % above 2Nx - 1 by the measured cost below.
#pragma once

namespace tdda {

// x-pass time (us) per length on MI355X: a 16 x 32 x 12 box with 8
// right-hand sides (so the loads and stores are the same for every length),
// complex64 and complex128, the kernel the planner picks for that length
// (register-resident, or the LDS transform where complex128 has no register
// split).  Measured with topo-dda-forwin64_amd/profile_lengths.py.
struct XLenCost {
    int L;
    float c64, c128;
};
constexpr XLenCost kXLenCost[] = {
    {32, 7.28f, 9.06f}, {36, 7.28f, 9.43f}, {40, 7.92f, 10.27f}, {45, 8.44f, 11.18f},
    {48, 8.15f, 10.70f}, {50, 8.42f, 11.00f}, {54, 8.66f, 11.71f}, {60, 9.19f, 12.42f},
    {64, 9.39f, 14.82f}, {72, 10.25f, 15.55f}, {75, 12.05f, 19.63f}, {80, 10.23f, 15.62f},
    {81, 10.62f, 15.05f}, {90, 11.14f, 16.01f}, {96, 12.52f, 21.00f}, {100, 11.55f, 16.77f},
    {108, 12.77f, 21.45f}, {120, 15.20f, 24.01f}, {125, 21.70f, 53.75f}, {128, 14.17f, 22.62f},
    {135, 16.36f, 27.50f}, {144, 15.69f, 25.55f}, {150, 17.45f, 28.62f}, {160, 17.60f, 32.62f},
    {162, 16.84f, 83.13f}, {180, 21.41f, 35.35f}, {192, 20.29f, 34.45f}, {200, 18.90f, 83.44f},
    {216, 24.37f, 99.23f}, {225, 26.10f, 48.62f}, {240, 26.44f, 48.87f}, {243, 34.65f, 124.01f},
    {250, 29.06f, 108.10f}, {256, 25.75f, 47.76f}, {270, 32.79f, 157.81f}, {288, 33.13f, 145.04f},
    {300, 33.81f, 143.55f}, {320, 33.85f, 108.82f}, {324, 36.03f, 150.53f}, {360, 37.50f, 155.14f},
    {375, 49.43f, 191.13f}, {384, 48.36f, 135.21f}, {400, 39.82f, 192.03f}, {405, 52.14f, 257.00f},
    {432, 51.34f, 211.12f}, {450, 54.61f, 254.48f}, {480, 75.61f, 194.73f}, {486, 56.03f, 313.41f},
    {500, 58.53f, 226.83f}, {512, 59.72f, 151.80f}, {540, 79.82f, 481.70f},
};

inline int grid_next_smooth(int n)
{
    for (n = n < 1 ? 1 : n;; n++) {
        int m = n;
        for (int p : {2, 3, 5})
            while (m % p == 0) m /= p;
        if (m == 1) return n;
    }
}

// prec: the operator's precision bits (TDDA_C128 = 1, TDDA_C64 = 2).
inline int default_grid_len(int axis, int n, int prec)
{
    int L = grid_next_smooth(n);
    if (axis == 1) {
        // the y lengths 75 and 125 (radix-5 splits) are slower than the next
        // smooth length over a whole application, whose bytes are 7 % and 2 %
        // more (measured: 125 -> 128: complex64 74.0 -> 70.0 us, complex128
        // 101.5 -> 90.8 us on 32 x 63 x 12 with 8 right-hand sides)
        if (L == 75 || L == 125) L = grid_next_smooth(L + 1);
        return L;
    }
    if (axis != 0) return L;
    // candidates: the table's lengths within 30 % above n; with both
    // precisions built, the one whose worse precision is least slower than
    // that precision's best candidate (ties: the smaller total)
    const double cap = 1.3 * n + 1;
    double m64 = 0, m128 = 0;
    for (const XLenCost& e : kXLenCost) {
        if (e.L < n || e.L > cap) continue;
        if (m64 == 0 || e.c64 < m64) m64 = e.c64;
        if (m128 == 0 || e.c128 < m128) m128 = e.c128;
    }
    int best = 0;
    double bk = 0, bs = 0;
    for (const XLenCost& e : kXLenCost) {
        if (e.L < n || e.L > cap) continue;
        const double r64 = (prec & 2) ? e.c64 / m64 : 0.0, r128 = (prec & 1) ? e.c128 / m128 : 0.0;
        const double k = r64 > r128 ? r64 : r128;
        const double sum = ((prec & 2) ? e.c64 : 0.0) + ((prec & 1) ? e.c128 : 0.0);
        if (best == 0 || k < bk - 1e-9 || (k <= bk + 1e-9 && sum < bs)) {
            best = e.L;
            bk = k;
            bs = sum;
        }
    }
    return best ? best : L;
}

}  // namespace tdda
