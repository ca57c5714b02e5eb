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
// (register-resident up to 540 in both precisions, above 540 in complex64
// only: complex128 runs the LDS transform there).  Measured with topo-dda-forwin64_amd/profile_lengths.py.
struct XLenCost {
    int L;
    float c64, c128;
};
constexpr XLenCost kXLenCost[] = {
    {32, 7.32f, 9.09f}, {36, 7.31f, 9.47f}, {40, 7.96f, 10.30f}, {45, 8.47f, 11.23f},
    {48, 8.17f, 10.75f}, {50, 8.45f, 11.02f}, {54, 8.69f, 11.71f}, {60, 9.24f, 12.54f},
    {64, 9.45f, 14.96f}, {72, 10.33f, 15.60f}, {75, 12.06f, 19.52f}, {80, 10.27f, 15.68f},
    {81, 10.66f, 15.01f}, {90, 11.19f, 15.94f}, {96, 12.58f, 21.15f}, {100, 11.62f, 16.70f},
    {108, 12.83f, 21.56f}, {120, 15.24f, 23.98f}, {125, 21.80f, 51.81f}, {128, 14.22f, 22.77f},
    {135, 16.41f, 27.49f}, {144, 15.75f, 25.45f}, {150, 17.52f, 28.66f}, {160, 17.67f, 32.95f},
    {162, 16.90f, 31.42f}, {180, 21.47f, 35.42f}, {192, 20.36f, 34.49f}, {200, 18.98f, 35.48f},
    {216, 24.54f, 49.96f}, {225, 26.22f, 48.91f}, {240, 26.48f, 49.11f}, {243, 34.53f, 68.50f},
    {250, 29.21f, 68.18f}, {256, 26.61f, 48.77f}, {270, 32.95f, 59.30f}, {288, 33.21f, 59.63f},
    {300, 33.65f, 59.38f}, {320, 33.85f, 61.11f}, {324, 36.04f, 66.07f}, {360, 37.28f, 68.84f},
    {375, 49.22f, 92.88f}, {384, 48.52f, 87.77f}, {400, 39.96f, 72.84f}, {405, 52.44f, 98.44f},
    {432, 51.66f, 93.75f}, {450, 54.98f, 101.85f}, {480, 75.92f, 97.41f}, {486, 55.35f, 106.11f},
    {500, 58.63f, 104.12f}, {512, 60.22f, 109.76f}, {540, 80.12f, 169.05f}, {576, 83.16f, 399.07f},
    {600, 85.83f, 420.16f}, {625, 89.91f, 433.97f}, {640, 104.62f, 410.90f}, {648, 90.35f, 512.22f},
    {675, 93.46f, 518.06f}, {720, 114.11f, 551.53f}, {729, 98.42f, 612.23f}, {750, 114.18f, 536.81f},
    {768, 112.50f, 480.86f}, {800, 117.26f, 752.64f}, {810, 126.21f, 1101.62f}, {864, 125.62f, 991.51f},
    {900, 134.46f, 1002.25f}, {960, 134.89f, 857.64f}, {972, 153.72f, 1158.07f}, {1000, 157.67f, 876.11f},
    {1024, 134.79f, 891.96f},
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
