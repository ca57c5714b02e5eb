// tdda_grid.hpp -- default zero-padded FFT grid of an operator (host only).
//
// Any M >= 2N - 1 along an axis gives the same linear convolution.  The y and
// z lengths set the bytes of every pass (the intermediates hold My Mz
// points per x row), so they stay the smallest 2,3,5-smooth length, except
// where that length is known slow.  The x length sets only arithmetic: the
// x pass loads and stores the Nx rows that exist and transforms them in
// registers or LDS, so its cost is that of the transform, which varies with
// the factors of the length more than with the length itself (125 = 5^3
// takes 1.5x the time of 128 in complex64 and 2.3x in complex128).  The default x length is therefore chosen
// among the lengths within 30 % above 2Nx - 1 by the measured cost below.
#pragma once

namespace tdda {

// x-pass time (us) per length on MI355X: a 16 x 32 x 12 box with 8
// right-hand sides (so the loads and stores are the same for every length),
// complex64 and complex128, the kernel the planner picks for that length
// (register-resident up to 540 in both precisions).  Measured with topo-dda-forwin64_amd/profile_lengths.py.
struct XLenCost {
    int L;
    float c64, c128;
};
constexpr XLenCost kXLenCost[] = {
    {32, 7.31f, 9.07f}, {36, 7.30f, 9.48f}, {40, 7.95f, 10.30f}, {45, 8.50f, 11.20f},
    {48, 8.17f, 10.74f}, {50, 8.44f, 11.02f}, {54, 8.69f, 11.71f}, {60, 9.25f, 12.52f},
    {64, 9.46f, 14.86f}, {72, 10.27f, 15.59f}, {75, 12.04f, 19.60f}, {80, 10.26f, 15.64f},
    {81, 10.69f, 15.09f}, {90, 11.18f, 16.00f}, {96, 12.55f, 21.12f}, {100, 11.60f, 16.86f},
    {108, 12.78f, 21.63f}, {120, 15.22f, 24.13f}, {125, 21.80f, 51.98f}, {128, 14.18f, 22.77f},
    {135, 16.41f, 27.73f}, {144, 15.75f, 25.66f}, {150, 17.51f, 28.87f}, {160, 17.65f, 33.01f},
    {162, 16.86f, 31.88f}, {180, 21.49f, 35.89f}, {192, 20.36f, 35.00f}, {200, 18.94f, 36.21f},
    {216, 24.49f, 50.28f}, {225, 26.17f, 49.30f}, {240, 26.52f, 49.67f}, {243, 34.65f, 69.24f},
    {250, 29.15f, 67.75f}, {256, 26.51f, 48.71f}, {270, 33.02f, 59.24f}, {288, 33.27f, 60.62f},
    {300, 33.81f, 61.08f}, {320, 33.97f, 61.49f}, {324, 36.22f, 66.39f}, {360, 37.59f, 70.01f},
    {375, 49.61f, 93.75f}, {384, 48.97f, 88.75f}, {400, 40.14f, 72.85f}, {405, 52.72f, 99.05f},
    {432, 52.15f, 95.79f}, {450, 55.64f, 103.56f}, {480, 76.19f, 98.39f}, {486, 56.50f, 107.33f},
    {500, 59.16f, 105.74f}, {512, 60.66f, 110.46f}, {540, 81.24f, 169.49f},
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
    auto score = [&](const XLenCost& e) {
        const double r64 = (prec & 2) ? e.c64 / m64 : 0.0, r128 = (prec & 1) ? e.c128 / m128 : 0.0;
        return r64 > r128 ? r64 : r128;
    };
    int best = 0;
    double bk = 0, bs = 0, k0 = 0;
    for (const XLenCost& e : kXLenCost) {
        if (e.L < n || e.L > cap) continue;
        const double k = score(e);
        const double sum = ((prec & 2) ? e.c64 : 0.0) + ((prec & 1) ? e.c128 : 0.0);
        if (e.L == L) k0 = k;
        if (best == 0 || k < bk - 1e-9 || (k <= bk + 1e-9 && sum < bs)) {
            best = e.L;
            bk = k;
            bs = sum;
        }
    }
    // the smallest smooth length stays unless another is more than 5 %
    // better: the table is measured on one box shape with the planner's
    // tiles, and closer calls can go either way on others
    if (k0 > 0 && bk * 1.05 >= k0) return L;
    return best ? best : L;
}

}  // namespace tdda
