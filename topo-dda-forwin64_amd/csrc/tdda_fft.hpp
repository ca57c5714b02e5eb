// tdda_fft.hpp -- complex arithmetic and the LDS-staged mixed-radix FFT used by
// every pass of the DDA operator.
//
// The 3D transforms of the operator are split into one pass per axis.  A pass
// stages a tile of pencils in LDS, runs self-sorting (Stockham) radix stages
// over them and writes the tile back.  Radices 2, 3, 4, 5 and 8 are supported,
// so any FFT length whose prime factors are 2, 3 and 5 can be planned.
//
// Twiddles come from tables of the L-th roots of unity computed on the host in
// long double (forward sign, exp(-2 pi i q / L)); the inverse uses conjugates.
#pragma once
#include <hip/hip_runtime.h>
#include <type_traits>

namespace tdda {

template <typename T>
struct alignas(2 * sizeof(T)) Cx {
    T x, y;
};

template <typename T>
__device__ __forceinline__ Cx<T> cadd(Cx<T> a, Cx<T> b) { return {a.x + b.x, a.y + b.y}; }
template <typename T>
__device__ __forceinline__ Cx<T> csub(Cx<T> a, Cx<T> b) { return {a.x - b.x, a.y - b.y}; }
template <typename T>
__device__ __forceinline__ Cx<T> cmul(Cx<T> a, Cx<T> b)
{
    return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
// a * conj(b)
template <typename T>
__device__ __forceinline__ Cx<T> cmulc(Cx<T> a, Cx<T> b)
{
    return {a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y};
}
template <typename T>
__device__ __forceinline__ Cx<T> cscale(Cx<T> a, T s) { return {a.x * s, a.y * s}; }
template <typename T>
__device__ __forceinline__ Cx<T> cneg(Cx<T> a) { return {-a.x, -a.y}; }
// multiply by -i (forward) or +i (inverse)
template <typename T, bool INV>
__device__ __forceinline__ Cx<T> rot90(Cx<T> a)
{
    if (INV) return {-a.y, a.x};
    return {a.y, -a.x};
}

// Threads per workgroup of every operator pass.
constexpr int kThreadsFft = 256;

// Plan of one 1D transform: length and its radix sequence.
// tb: base of the two-level twiddle table, exp(-2 pi i q / L) =
// T0[q % tb] * T1[q / tb] with T0 of tb entries and T1 of ceil(L / tb).
struct FftPlan {
    int L;
    int nst;
    int rad[12];
    int tb;
};

// ---------------------------------------------------------------------------
// Small DFTs in registers.  Forward: y_q = sum_m v_m exp(-2 pi i m q / R).
// ---------------------------------------------------------------------------
template <typename T, bool INV>
__device__ __forceinline__ void dft2(Cx<T>* v)
{
    Cx<T> a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
}

template <typename T, bool INV>
__device__ __forceinline__ void dft4(Cx<T>* v)
{
    Cx<T> t0 = cadd(v[0], v[2]);
    Cx<T> t1 = csub(v[0], v[2]);
    Cx<T> t2 = cadd(v[1], v[3]);
    Cx<T> t3 = rot90<T, INV>(csub(v[1], v[3]));
    v[0] = cadd(t0, t2);
    v[2] = csub(t0, t2);
    v[1] = cadd(t1, t3);
    v[3] = csub(t1, t3);
}

template <typename T, bool INV>
__device__ __forceinline__ void dft8(Cx<T>* v)
{
    const T h = T(0.70710678118654752440084436210484903928);
    Cx<T> a[4], b[4];
#pragma unroll
    for (int m = 0; m < 4; m++) {
        a[m] = cadd(v[m], v[m + 4]);
        b[m] = csub(v[m], v[m + 4]);
    }
    // b[m] *= W8^m, W8 = exp(-+ i pi / 4)
    b[1] = INV ? Cx<T>{h * (b[1].x - b[1].y), h * (b[1].x + b[1].y)}
               : Cx<T>{h * (b[1].x + b[1].y), h * (b[1].y - b[1].x)};
    b[2] = rot90<T, INV>(b[2]);
    b[3] = INV ? Cx<T>{-h * (b[3].x + b[3].y), h * (b[3].x - b[3].y)}
               : Cx<T>{h * (b[3].y - b[3].x), -h * (b[3].x + b[3].y)};
    dft4<T, INV>(a);
    dft4<T, INV>(b);
#pragma unroll
    for (int q = 0; q < 4; q++) {
        v[2 * q] = a[q];
        v[2 * q + 1] = b[q];
    }
}

template <typename T, bool INV>
__device__ __forceinline__ void dft3(Cx<T>* v)
{
    const T c = T(-0.5);
    const T s = T(0.86602540378443864676372317075293618347);
    Cx<T> t1 = cadd(v[1], v[2]);
    Cx<T> y0 = cadd(v[0], t1);
    Cx<T> m1 = {v[0].x + c * t1.x, v[0].y + c * t1.y};
    Cx<T> d = csub(v[1], v[2]);
    Cx<T> m2 = rot90<T, INV>(Cx<T>{s * d.x, s * d.y});
    v[0] = y0;
    v[1] = cadd(m1, m2);
    v[2] = csub(m1, m2);
}

template <typename T, bool INV>
__device__ __forceinline__ void dft5(Cx<T>* v)
{
    const T c1 = T(0.30901699437494742410229341718281905886);
    const T c2 = T(-0.80901699437494742410229341718281905886);
    const T s1 = T(0.95105651629515357211643933337938214340);
    const T s2 = T(0.58778525229247312916870595463907276860);
    Cx<T> t1 = cadd(v[1], v[4]);
    Cx<T> t2 = cadd(v[2], v[3]);
    Cx<T> t3 = csub(v[1], v[4]);
    Cx<T> t4 = csub(v[2], v[3]);
    Cx<T> y0 = cadd(v[0], cadd(t1, t2));
    Cx<T> a1 = {v[0].x + c1 * t1.x + c2 * t2.x, v[0].y + c1 * t1.y + c2 * t2.y};
    Cx<T> a2 = {v[0].x + c2 * t1.x + c1 * t2.x, v[0].y + c2 * t1.y + c1 * t2.y};
    Cx<T> b1 = rot90<T, INV>(Cx<T>{s1 * t3.x + s2 * t4.x, s1 * t3.y + s2 * t4.y});
    Cx<T> b2 = rot90<T, INV>(Cx<T>{s2 * t3.x - s1 * t4.x, s2 * t3.y - s1 * t4.y});
    v[0] = y0;
    v[1] = cadd(a1, b1);
    v[4] = csub(a1, b1);
    v[2] = cadd(a2, b2);
    v[3] = csub(a2, b2);
}

template <typename T, int R, bool INV>
__device__ __forceinline__ void dft(Cx<T>* v)
{
    if constexpr (R == 2) dft2<T, INV>(v);
    else if constexpr (R == 3) dft3<T, INV>(v);
    else if constexpr (R == 4) dft4<T, INV>(v);
    else if constexpr (R == 5) dft5<T, INV>(v);
    else dft8<T, INV>(v);
}

// ---------------------------------------------------------------------------
// Small DFTs of any length in {2, 3, 4, 5, 6, 8, 9, 10, 12, 15, 16, 18, 20,
// 24, 25, 27, 30, 32, 36, 40}, held entirely in registers, for the register-resident
// x and y passes.  Lengths other
// than 2, 3, 4, 5, 8 are split as R = A * B (decimation in time:
// n = B a + b, k = k1 + A k2):
//   X[k1 + A k2] = sum_b w_B^(b k2) w_R^(b k1) sum_a v[B a + b] w_A^(a k1)
// with the twiddles w_R^(b k1) as compile-time constants.
// ---------------------------------------------------------------------------
struct CosSin {
    double c, s;
};

// cos and sin of x for |x| <= pi/4 (Taylor series, 13 terms: below one ulp)
constexpr CosSin ct_cs_small(double x)
{
    const double x2 = x * x;
    double c = 1.0, s = x, tc = 1.0, ts = x;
    for (int k = 1; k <= 13; k++) {
        tc *= -x2 / (double)((2 * k - 1) * (2 * k));
        ts *= -x2 / (double)((2 * k) * (2 * k + 1));
        c += tc;
        s += ts;
    }
    return CosSin{c, s};
}

// (cos, sin) of 2 pi m / R at compile time: exact reduction to a quadrant
// q and a remainder angle (pi/2) r / R, then to [0, pi/4] by the complement.
constexpr CosSin ct_root(int m, int R)
{
    m %= R;
    if (m < 0) m += R;
    const int q = (4 * m) / R, r = 4 * m - q * R;
    const double hp = 1.57079632679489661923132169163975144;
    CosSin b{1.0, 0.0};
    if (2 * r <= R) {
        b = ct_cs_small(hp * (double)r / (double)R);
    } else {
        const CosSin t = ct_cs_small(hp * (double)(R - r) / (double)R);
        b = CosSin{t.s, t.c};
    }
    switch (q & 3) {
    case 0: return b;
    case 1: return CosSin{-b.s, b.c};
    case 2: return CosSin{-b.c, -b.s};
    default: return CosSin{b.s, -b.c};
    }
}

constexpr bool sdft_ok(int R)
{
    return R == 2 || R == 3 || R == 4 || R == 5 || R == 6 || R == 8 || R == 9 || R == 10 || R == 12 || R == 15 ||
           R == 16 || R == 18 || R == 20 || R == 24 || R == 25 || R == 27 || R == 30 || R == 32 || R == 36 ||
           R == 40;
}
// first factor A of the split R = A * B
constexpr int sdft_split(int R) { return (R % 4 == 0) ? 4 : (R % 2 == 0) ? 2 : (R % 3 == 0) ? 3 : 5; }

// Calls f(std::integral_constant<int, I>) for I = I0 .. N-1 (unrolled at
// compile time, so every index is a constant).
template <int I, int N, typename F>
__device__ __forceinline__ void sfor(F&& f)
{
    if constexpr (I < N) {
        f(std::integral_constant<int, I>{});
        sfor<I + 1, N>(f);
    }
}

template <typename T, int R, bool INV>
__device__ __forceinline__ void sdft(Cx<T>* v)
{
    static_assert(R == 1 || sdft_ok(R), "unsupported register DFT length");
    if constexpr (R == 1) {
        return;
    } else if constexpr (R == 2 || R == 3 || R == 4 || R == 5 || R == 8) {
        dft<T, R, INV>(v);
    } else {
        constexpr int A = sdft_split(R), B = R / A;
        Cx<T> u[B][A];
        sfor<0, B>([&](auto bc) {
            constexpr int b = decltype(bc)::value;
            Cx<T> t[A];
            sfor<0, A>([&](auto ac) {
                constexpr int a = decltype(ac)::value;
                t[a] = v[B * a + b];
            });
            sdft<T, A, INV>(t);
            sfor<0, A>([&](auto kc) {
                constexpr int k1 = decltype(kc)::value;
                if constexpr (b * k1 == 0) {
                    u[b][k1] = t[k1];
                } else {
                    // forward: times exp(-2 pi i b k1 / R); inverse: its conjugate
                    constexpr CosSin w = ct_root(b * k1, R);
                    const T c = (T)w.c, s = INV ? (T)w.s : (T)(-w.s);
                    u[b][k1] = Cx<T>{t[k1].x * c - t[k1].y * s, t[k1].x * s + t[k1].y * c};
                }
            });
        });
        sfor<0, A>([&](auto kc) {
            constexpr int k1 = decltype(kc)::value;
            Cx<T> t[B];
            sfor<0, B>([&](auto bc) {
                constexpr int b = decltype(bc)::value;
                t[b] = u[b][k1];
            });
            sdft<T, B, INV>(t);
            sfor<0, B>([&](auto k2c) {
                constexpr int k2 = decltype(k2c)::value;
                v[k1 + A * k2] = t[k2];
            });
        });
    }
}

// Factors L = L1 * L2 of the register-resident x pass: both in the small-DFT
// set, L1 >= L2 and as close as possible (L1 = 0: no such split).  The
// larger factor goes to the step that holds the Green product: the steps
// with the global loads and stores then hold fewer values (160 = 16 x 10
// needs 146 VGPRs in complex128, 10 x 16 needs 171).
constexpr int pxr_l1(int L)
{
    int best = 0;
    for (int a = 2; a * a <= L; a++)
        if (L % a == 0 && sdft_ok(a) && sdft_ok(L / a)) best = L / a;
    return best;
}

// One Stockham stage over NP pencils of length L held in LDS (pencil stride
// ps).  Ns is the length of the sub-transforms already completed.
template <typename T, int R, bool INV>
__device__ __forceinline__ void stockham_stage(const Cx<T>* __restrict__ src, Cx<T>* __restrict__ dst,
                                               int NP, int ps, int L, int Ns, const Cx<T>* __restrict__ tw)
{
    const int LR = L / R;
    const int total = NP * LR;
    const int twstep = L / (Ns * R);
    for (int b = threadIdx.x; b < total; b += blockDim.x) {
        const int p = b / LR;
        const int j = b - p * LR;
        const int k = j % Ns;
        const Cx<T>* s = src + p * ps + j;
        Cx<T> v[R];
#pragma unroll
        for (int m = 0; m < R; m++) v[m] = s[m * LR];
        if (Ns > 1) {
#pragma unroll
            for (int m = 1; m < R; m++) {
                const Cx<T> w = tw[m * k * twstep];
                v[m] = INV ? cmulc(v[m], w) : cmul(v[m], w);
            }
        }
        dft<T, R, INV>(v);
        Cx<T>* d = dst + p * ps + (j - k) * R + k;
#pragma unroll
        for (int m = 0; m < R; m++) d[m * Ns] = v[m];
    }
}

// Full transform of NP pencils: data starts in a, b is scratch; returns the
// buffer that holds the result.  Ends with a barrier.
template <typename T, bool INV>
__device__ Cx<T>* lds_fft(Cx<T>* a, Cx<T>* b, int NP, int ps, const FftPlan& pl, const Cx<T>* __restrict__ tw)
{
    int Ns = 1;
    const int L = pl.L;
    const int nst = pl.nst;
    for (int s = 0; s < nst; s++) {
        const int R = pl.rad[s];
        switch (R) {
        case 8: stockham_stage<T, 8, INV>(a, b, NP, ps, L, Ns, tw); break;
        case 4: stockham_stage<T, 4, INV>(a, b, NP, ps, L, Ns, tw); break;
        case 2: stockham_stage<T, 2, INV>(a, b, NP, ps, L, Ns, tw); break;
        case 3: stockham_stage<T, 3, INV>(a, b, NP, ps, L, Ns, tw); break;
        default: stockham_stage<T, 5, INV>(a, b, NP, ps, L, Ns, tw); break;
        }
        __syncthreads();
        Cx<T>* t = a;
        a = b;
        b = t;
        Ns *= R;
    }
    return a;
}

}  // namespace tdda

namespace tdda {

// ---------------------------------------------------------------------------
// LDS transform used by the operator passes.
//
// Pencil p, element i lives at p * ps + lpad<T>(i): spare slots spread the
// strided Stockham writes of the early stages over the LDS banks.  The
// runtime-plan transform (pp_fft) takes its twiddles from an LDS copy of a
// two-level table of the L-th roots of unity (tb + ceil(L / tb) entries,
// tb ~ sqrt(L)): one table product per butterfly, then powers by
// recurrence; its index arithmetic uses float reciprocals instead of integer
// division.  The compile-time transforms (ct_fft) read the full table.
// ---------------------------------------------------------------------------
// One spare slot every 2^shift elements: 16 for complex128 (16-byte
// elements, whose strided first-stage writes conflict most), 32 for
// complex64 (measured: denser padding costs complex64 more in occupancy than
// it saves in conflicts).
#ifndef TDDA_LDS_PAD_SHIFT_C128
#define TDDA_LDS_PAD_SHIFT_C128 4
#endif
#ifndef TDDA_LDS_PAD_SHIFT_C64
#define TDDA_LDS_PAD_SHIFT_C64 5
#endif
template <typename T>
__host__ __device__ constexpr int pad_shift()
{
    return sizeof(T) == 8 ? TDDA_LDS_PAD_SHIFT_C128 : TDDA_LDS_PAD_SHIFT_C64;
}
template <typename T>
__device__ __forceinline__ int lpad(int i)
{
    return i + (i >> pad_shift<T>());
}
template <typename T>
__host__ __device__ __forceinline__ int lds_pencil_stride(int L)
{
    return L + (L >> pad_shift<T>()) + 1;
}

// q = n / d, r = n % d for 0 <= n < 2^24 via a float reciprocal and one
// correction step (exact in that range).  Branch-free, with a 24-bit
// multiply (full rate, unlike the 32-bit one).
__device__ __forceinline__ void fdivmod(int n, int d, float rd, int& q, int& r)
{
    q = (int)((float)n * rd);
    r = n - (int)__umul24((unsigned)q, (unsigned)d);
    const bool lo = r < 0, hi = r >= d;
    q = lo ? q - 1 : (hi ? q + 1 : q);
    r = lo ? r + d : (hi ? r - d : r);
}

// One Stockham stage of the operator passes: NP pencils, src -> dst (LDS,
// padded layout), twiddles from the two-level LDS table by recurrence.
template <typename T, int R, bool INV>
__device__ __forceinline__ void pp_stage(const Cx<T>* __restrict__ src, Cx<T>* __restrict__ dst, int NP, int ps,
                                         int L, int Ns, const Cx<T>* __restrict__ tw, int tb)
{
    const int LR = L / R;
    const int total = NP * LR;
    const int twstep = L / (Ns * R);
    const float rLR = 1.0f / (float)LR;
    const float rNs = 1.0f / (float)Ns;
    const float rtb = 1.0f / (float)tb;
    for (int b = threadIdx.x; b < total; b += kThreadsFft) {
        int p, j, q, kk;
        fdivmod(b, LR, rLR, p, j);
        fdivmod(j, Ns, rNs, q, kk);
        const Cx<T>* sp = src + p * ps;
        Cx<T> v[R];
#pragma unroll
        for (int m = 0; m < R; m++) v[m] = sp[lpad<T>(j + m * LR)];
        if (Ns > 1) {
            // w^m by recurrence from w = exp(-+2 pi i kk twstep / L)
            int q1, q0;
            fdivmod(kk * twstep, tb, rtb, q1, q0);
            const Cx<T> w1 = cmul(tw[q0], tw[tb + q1]);
            Cx<T> w = w1;
#pragma unroll
            for (int m = 1; m < R; m++) {
                v[m] = INV ? cmulc(v[m], w) : cmul(v[m], w);
                if (m + 1 < R) w = cmul(w, w1);
            }
        }
        dft<T, R, INV>(v);
        Cx<T>* dp = dst + p * ps;
        const int base = (j - kk) * R + kk;
#pragma unroll
        for (int m = 0; m < R; m++) dp[lpad<T>(base + m * Ns)] = v[m];
    }
}

// Transform of NP pencils held in a (padded layout); b is scratch of the
// same size.  Returns the buffer holding the result; ends with a barrier.
template <typename T, bool INV>
__device__ __forceinline__ Cx<T>* pp_fft(Cx<T>* a, Cx<T>* b, int NP, int ps, const FftPlan& pl, const Cx<T>* tw)
{
    int Ns = 1;
    const int L = pl.L;
    for (int s = 0; s < pl.nst; s++) {
        const int R = pl.rad[s];
        switch (R) {
        case 8: pp_stage<T, 8, INV>(a, b, NP, ps, L, Ns, tw, pl.tb); break;
        case 4: pp_stage<T, 4, INV>(a, b, NP, ps, L, Ns, tw, pl.tb); break;
        case 2: pp_stage<T, 2, INV>(a, b, NP, ps, L, Ns, tw, pl.tb); break;
        case 3: pp_stage<T, 3, INV>(a, b, NP, ps, L, Ns, tw, pl.tb); break;
        default: pp_stage<T, 5, INV>(a, b, NP, ps, L, Ns, tw, pl.tb); break;
        }
        __syncthreads();
        Cx<T>* t = a;
        a = b;
        b = t;
        Ns *= R;
    }
    return a;
}

// Entries of the two-level twiddle table of a plan.
__host__ __device__ constexpr int tw2_size(int L, int tb) { return tb + (L + tb - 1) / tb; }
// LDS slots reserved for it (even, so the data after it stays 16-byte aligned).
__host__ __device__ constexpr int tw2_slots(int L, int tb) { return (tw2_size(L, tb) + 1) & ~1; }

// Cooperative copy of the two-level twiddle table into LDS.
template <typename T>
__device__ __forceinline__ void load_twiddles(Cx<T>* __restrict__ dst, const Cx<T>* __restrict__ src,
                                              const FftPlan& pl)
{
    const int n = tw2_size(pl.L, pl.tb);
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        using V = typename std::conditional<sizeof(T) == 4, float __attribute__((ext_vector_type(2))),
                                            double __attribute__((ext_vector_type(2)))>::type;
        const V v = *(const __attribute__((address_space(1))) V*)(src + i);
        dst[i] = Cx<T>{v.x, v.y};
    }
}

}  // namespace tdda

namespace tdda {

// ---------------------------------------------------------------------------
// Compile-time specialisation of the pass transform for a fixed length L:
// radix sequence, strides, twiddle steps and pencil stride are constants, so
// every index division becomes a multiply-shift and the stage loops unroll.
// Same data layout as pp_fft; twiddles from the full table of L roots, or
// from the two-level table where the full one would cost occupancy.
// ---------------------------------------------------------------------------
constexpr int ct_nst(int L)
{
    int n = L, e2 = 0, c = 0;
    while (n % 2 == 0) {
        n /= 2;
        e2++;
    }
    c = e2 / 3 + (e2 % 3 ? 1 : 0);
    while (n % 3 == 0) {
        n /= 3;
        c++;
    }
    while (n % 5 == 0) {
        n /= 5;
        c++;
    }
    return c;
}

constexpr int ct_rad(int L, int s)
{
    int n = L, e2 = 0, k = 0;
    while (n % 2 == 0) {
        n /= 2;
        e2++;
    }
    while (e2 >= 3) {
        if (k++ == s) return 8;
        e2 -= 3;
    }
    if (e2 == 2 && k++ == s) return 4;
    if (e2 == 1 && k++ == s) return 2;
    while (n % 3 == 0) {
        n /= 3;
        if (k++ == s) return 3;
    }
    while (n % 5 == 0) {
        n /= 5;
        if (k++ == s) return 5;
    }
    return 1;
}

// base of the two-level twiddle table (make_plan: smallest tb, tb^2 >= L)
__host__ __device__ constexpr int ct_tb(int L)
{
    int t = 1;
    while (t * t < L) t++;
    return t;
}

constexpr int kPruneIn = 1, kPruneOut = 2;
// Entries a pruned first stage still reads: [0, ceil(R0/2) L/R0).
constexpr int ct_zin_limit(int L)
{
    return L > 1 ? ((ct_rad(L, 0) + 1) / 2) * (L / ct_rad(L, 0)) : L;
}

// Butterfly groups per lane that a stage interleaves (unroll factor of
// the per-lane butterfly loop); full unrolling spills in the x pass.
#ifndef TDDA_STAGE_UNROLL
#define TDDA_STAGE_UNROLL 1
#endif

// Padded LDS offset of element i + m * STEP.  When STEP is a multiple of the
// padding group, lpad(i + m STEP) = lpad(i) + m (STEP + STEP / group), so
// the R accesses of a butterfly share one address and differ by constants.
template <typename T, unsigned STEP>
__device__ __forceinline__ unsigned lpad_step(unsigned i, int m)
{
    constexpr int SH = pad_shift<T>();
    if constexpr (STEP % (1u << SH) == 0) return (i + (i >> SH)) + (unsigned)m * (STEP + (STEP >> SH));
    else {
        const unsigned x = i + (unsigned)m * STEP;
        return x + (x >> SH);
    }
}

// Twiddles of a butterfly of a compile-time stage from the full LDS table
// tw[q] = exp(-2 pi i q / L): v[m] *= w^m with w = tw[q], read directly for
// m = 1, 2 (and 4), the other powers as one product of two table entries.
template <typename T, int R, bool INV>
__device__ __forceinline__ void twiddle_direct(Cx<T>* v, const Cx<T>* __restrict__ tw, unsigned q)
{
    auto app = [&](int m, Cx<T> w) { v[m] = INV ? cmulc(v[m], w) : cmul(v[m], w); };
    const Cx<T> w1 = tw[q];
    app(1, w1);
    if constexpr (R >= 3) {
        const Cx<T> w2 = tw[2 * q];
        app(2, w2);
        if constexpr (R == 4) {
            app(3, cmul(w1, w2));
        } else if constexpr (R == 5) {
            app(3, cmul(w1, w2));
            app(4, cmul(w2, w2));
        } else if constexpr (R == 8) {
            const Cx<T> w4 = tw[4 * q];
            const Cx<T> w3 = cmul(w1, w2);
            app(3, w3);
            app(4, w4);
            app(5, cmul(w1, w4));
            app(6, cmul(w2, w4));
            app(7, cmul(w3, w4));
        }
    }
}

// LDS slots of a pass's twiddle table: the full table of L roots, or the
// two-level table (even, so the data after it stays 16-byte aligned).
__host__ __device__ constexpr int pass_tw_slots(bool full, int L, int tb)
{
    return full ? ((L + 1) & ~1) : tw2_slots(L, tb);
}

// Cooperative copy of a pass's twiddle table into LDS (see pass_tw_slots).
template <typename T, bool FULL>
__device__ __forceinline__ void load_pass_twiddles(Cx<T>* __restrict__ dst, const Cx<T>* __restrict__ two_level,
                                                   const Cx<T>* __restrict__ full, const FftPlan& pl)
{
    if constexpr (FULL) {
        using V = typename std::conditional<sizeof(T) == 4, float __attribute__((ext_vector_type(2))),
                                            double __attribute__((ext_vector_type(2)))>::type;
        for (int i = threadIdx.x; i < pl.L; i += blockDim.x) {
            const V w = *(const __attribute__((address_space(1))) V*)(full + i);
            dst[i] = Cx<T>{w.x, w.y};
        }
    } else {
        load_twiddles(dst, two_level, pl);
    }
}

// Wave-local stage: pencil p belongs to wavefront p % 4, so a stage only
// reads pencils its own wavefront wrote.  LDS instructions of one wavefront
// complete in order, so consecutive stages need no workgroup barrier, only
// a compiler-ordering barrier.
constexpr int kWaves = kThreadsFft / 64;

// Pruning of the zero-padded transforms (the passes keep L >= 2N - 1):
//   ZIN  -- input entries at or above N are zero: the first stage's inputs
//           j + m L/R with m >= ceil(R/2) lie at or above ceil(R/2) L/R >= N
//           and are taken as zero instead of read;
//   HOUT -- only outputs below N are needed: the last stage's outputs
//           base + m L/R with m >= ceil(R/2) are neither computed nor stored.
template <typename T, int L, int R, int NS, bool INV, int NPMAX, bool FULLTW, bool ZIN = false, bool HOUT = false>
__device__ __forceinline__ void ct_stage(const Cx<T>* __restrict__ src, Cx<T>* __restrict__ dst, int np,
                                         const Cx<T>* __restrict__ tw)
{
    constexpr int RH = (R + 1) / 2;
    constexpr unsigned LR = L / R;
    constexpr unsigned PS = L + (L >> pad_shift<T>()) + 1;
    constexpr unsigned TWSTEP = L / (NS * R);
    constexpr int NPW = (NPMAX + kWaves - 1) / kWaves;  // pencils per wavefront, at most
    constexpr int NBMAX = NPW * LR;
    const unsigned wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int npw = (np - (int)wave + kWaves - 1) / kWaves;  // this wavefront's pencils (>= 0)
    const int nb = npw * (int)LR;
#pragma unroll TDDA_STAGE_UNROLL
    for (int it = 0; it < (NBMAX + 63) / 64; it++) {
        const int b = (int)lane + it * 64;
        if (b < nb) {
            const unsigned pi = (unsigned)b / LR;
            const unsigned j = (unsigned)b - pi * LR;
            const unsigned kk = j % NS;
            const unsigned p = wave + pi * kWaves;
            const Cx<T>* sp = src + p * PS;
            Cx<T> v[R];
#pragma unroll
            for (int m = 0; m < R; m++)
                v[m] = (ZIN && m >= RH) ? Cx<T>{T(0), T(0)} : sp[lpad_step<T, LR>(j, m)];
            if constexpr (NS > 1) {
                if constexpr (FULLTW) {
                    twiddle_direct<T, R, INV>(v, tw, kk * TWSTEP);
                } else {
                    // two-level table, then powers by recurrence
                    constexpr unsigned TB = ct_tb(L);
                    const unsigned q = kk * TWSTEP;
                    const Cx<T> w1 = cmul(tw[q % TB], tw[TB + q / TB]);
                    Cx<T> w = w1;
#pragma unroll
                    for (int m = 1; m < R; m++) {
                        v[m] = INV ? cmulc(v[m], w) : cmul(v[m], w);
                        if (m + 1 < R) w = cmul(w, w1);
                    }
                }
            }
            dft<T, R, INV>(v);
            Cx<T>* dp = dst + p * PS;
            const unsigned base = (j - kk) * R + kk;
#pragma unroll
            for (int m = 0; m < R; m++)
                if (!(HOUT && m >= RH)) dp[lpad_step<T, NS>(base, m)] = v[m];
        }
    }
}

template <typename T, int L, int S, int NS, bool INV, int NPMAX, int PRUNE, bool FULLTW>
__device__ __forceinline__ Cx<T>* ct_stages(Cx<T>* a, Cx<T>* b, int np, const Cx<T>* tw)
{
    if constexpr (S == ct_nst(L)) {
        return a;
    } else {
        constexpr int R = ct_rad(L, S);
        constexpr bool zin = (PRUNE & kPruneIn) && S == 0;
        constexpr bool hout = (PRUNE & kPruneOut) && S + 1 == ct_nst(L);
        ct_stage<T, L, R, NS, INV, NPMAX, FULLTW, zin, hout>(a, b, np, tw);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        return ct_stages<T, L, S + 1, NS * R, INV, NPMAX, PRUNE, FULLTW>(b, a, np, tw);
    }
}

// In-place variant of ct_stage: a wavefront first reads every input of its
// butterflies into registers, then writes every output back into the same
// buffer.  LDS instructions of one wavefront complete in order, so no write
// can overtake a read of the same stage; the pencils of a wavefront are
// touched by no other wavefront.  Registers: NIT x R complex values per lane.
template <typename T, int L, int R, int NS, bool INV, int NPMAX, bool FULLTW, bool ZIN = false, bool HOUT = false>
__device__ __forceinline__ void ct_stage_ip(Cx<T>* buf, int np, const Cx<T>* __restrict__ tw)
{
    constexpr int RH = (R + 1) / 2;
    constexpr unsigned LR = L / R;
    constexpr unsigned PS = L + (L >> pad_shift<T>()) + 1;
    constexpr unsigned TWSTEP = L / (NS * R);
    constexpr int NPW = (NPMAX + kWaves - 1) / kWaves;
    // The last stage (NS R = L) writes each butterfly's outputs to the slots
    // it read, so its butterflies can run one at a time.
    constexpr bool LAST = NS * R == L;
    constexpr int NIT = LAST ? 1 : (NPW * (int)LR + 63) / 64;
    constexpr int NB = LAST ? (NPW * (int)LR + 63) / 64 : 1;
    const unsigned wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int npw = (np - (int)wave + kWaves - 1) / kWaves;
    const int nb = npw * (int)LR;
#pragma unroll 1
    for (int ob = 0; ob < NB; ob++) {
        Cx<T> v[NIT][R];
#pragma unroll
        for (int it = 0; it < NIT; it++) {
            const int b = (int)lane + (it + ob) * 64;
            if (b < nb) {
                const unsigned pi = (unsigned)b / LR;
                const unsigned j = (unsigned)b - pi * LR;
                const Cx<T>* sp = buf + (wave + pi * kWaves) * PS;
#pragma unroll
                for (int m = 0; m < R; m++)
                    v[it][m] = (ZIN && m >= RH) ? Cx<T>{T(0), T(0)} : sp[lpad_step<T, LR>(j, m)];
            }
        }
        if constexpr (!LAST) {
            // keep every read of this stage ahead of its writes
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }
#pragma unroll
        for (int it = 0; it < NIT; it++) {
            const int b = (int)lane + (it + ob) * 64;
            if (b < nb) {
                const unsigned pi = (unsigned)b / LR;
                const unsigned j = (unsigned)b - pi * LR;
                const unsigned kk = j % NS;
                if constexpr (NS > 1) {
                    if constexpr (FULLTW) {
                        twiddle_direct<T, R, INV>(v[it], tw, kk * TWSTEP);
                    } else {
                        constexpr unsigned TB = ct_tb(L);
                        const unsigned q = kk * TWSTEP;
                        const Cx<T> w1 = cmul(tw[q % TB], tw[TB + q / TB]);
                        Cx<T> w = w1;
#pragma unroll
                        for (int m = 1; m < R; m++) {
                            v[it][m] = INV ? cmulc(v[it][m], w) : cmul(v[it][m], w);
                            if (m + 1 < R) w = cmul(w, w1);
                        }
                    }
                }
                dft<T, R, INV>(v[it]);
                Cx<T>* dp = buf + (wave + pi * kWaves) * PS;
                const unsigned base = (j - kk) * R + kk;
#pragma unroll
                for (int m = 0; m < R; m++)
                    if (!(HOUT && m >= RH)) dp[lpad_step<T, NS>(base, m)] = v[it][m];
            }
        }
    }
}

template <typename T, int L, int S, int NS, bool INV, int NPMAX, int PRUNE, bool FULLTW>
__device__ __forceinline__ void ct_stages_ip(Cx<T>* a, int np, const Cx<T>* tw)
{
    if constexpr (S < ct_nst(L)) {
        constexpr int R = ct_rad(L, S);
        constexpr bool zin = (PRUNE & kPruneIn) && S == 0;
        constexpr bool hout = (PRUNE & kPruneOut) && S + 1 == ct_nst(L);
        ct_stage_ip<T, L, R, NS, INV, NPMAX, FULLTW, zin, hout>(a, np, tw);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        ct_stages_ip<T, L, S + 1, NS * R, INV, NPMAX, PRUNE, FULLTW>(a, np, tw);
    }
}

// Complex values per lane an in-place transform of up to NPMAX pencils of
// length L holds in registers (largest stage; the last stage runs one
// butterfly at a time).
constexpr int ct_ip_values(int L, int npmax)
{
    const int npw = (npmax + kWaves - 1) / kWaves;
    int most = 0;
    for (int s = 0; s < ct_nst(L); s++) {
        const int R = ct_rad(L, s);
        const int nit = s + 1 == ct_nst(L) ? 1 : (npw * (L / R) + 63) / 64;
        most = nit * R > most ? nit * R : most;
    }
    return most;
}

// In-place transform of np <= NPMAX pencils of the compile-time length L in
// a; same contract as ct_fft (returns a).
template <typename T, int L, bool INV, int NPMAX, int PRUNE = 0, bool FULLTW = false>
__device__ __forceinline__ Cx<T>* ct_fft_ip(Cx<T>* a, int np, const Cx<T>* tw)
{
    ct_stages_ip<T, L, 0, 1, INV, NPMAX, PRUNE, FULLTW>(a, np, tw);
    __syncthreads();
    return a;
}

// Transform of np <= NPMAX pencils of the compile-time length L; same
// contract as pp_fft (input visible to all threads, ends with a workgroup
// barrier).  PRUNE: kPruneIn (input zero at and above N; entries of
// [N, ct_zin_limit(L)) must still hold zeros) and/or kPruneOut (only
// outputs below N are valid afterwards).  FULLTW: tw is the full table of
// the L-th roots (else the two-level table).
template <typename T, int L, bool INV, int NPMAX, int PRUNE = 0, bool FULLTW = false>
__device__ __forceinline__ Cx<T>* ct_fft(Cx<T>* a, Cx<T>* b, int np, const Cx<T>* tw)
{
    Cx<T>* r = ct_stages<T, L, 0, 1, INV, NPMAX, PRUNE, FULLTW>(a, b, np, tw);
    __syncthreads();
    return r;
}

}  // namespace tdda
