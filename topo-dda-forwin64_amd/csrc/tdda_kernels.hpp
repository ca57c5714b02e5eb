// tdda_kernels.hpp -- device kernels of the DDA operator and solver.
//
// Operator (AProductCore::Aproduct, AProductCore.cpp:697-758, plus the
// al o b diagonal of DDAModel::Aproductwithalb, DDAModel.cpp:577-586):
//
//   y = IFFT( G^ . FFT(b) ) restricted to the lattice box, + al o b
//
// on a zero-padded grid Mx x My x Mz >= (2Nx-1, 2Ny-1, 2Nz-1).  The 3D
// transforms are pruned and split into five passes, each a kernel:
//
//   P1 z-forward  : box [Nx][Ny][Nz] (site vector)  -> w1 [Nx][Ny][Mz][3]
//   P2 y-forward  : w1                               -> w2 [Nx][My][Mz][3]
//   P3 x-fused    : w2 -> FFT_x, 6x3 symmetric Green product, IFFT_x -> w2
//   P4 y-inverse  : w2                               -> w1
//   P5 z-inverse  : w1 -> box, + diagonal, BiCGSTAB epilogues and reductions
//
// The Green spectrum is stored as one octant [Hx][Hy][Hz][6] (H = M/2 + 1):
// every tensor component is even or odd along each axis, so the other
// octants follow by sign flips.  The 1/M of the inverse transform is folded
// into the stored spectrum.
//
// Work is described by a table of jobs (one per structure x right-hand side)
// and each launch walks a flattened list of tiles over all jobs, so uneven
// structures share one launch.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "tdda_fft.hpp"
#include "tdda_tilemap.hpp"

namespace tdda {

constexpr int kThreads = kThreadsFft;

// Operator passes ask for at least TDDA_PASS_WAVES waves per SIMD so the
// compiler keeps their register footprint small enough for several
// workgroups per CU (latency hiding of the tile loads).
#ifndef TDDA_PASS_WAVES
#define TDDA_PASS_WAVES 4
#endif
#define TDDA_PASS_BOUNDS __launch_bounds__(kThreads, TDDA_PASS_WAVES)
// Register-resident kernels with a factor above 16 (complex64, lengths up to
// 1024): waves per SIMD they are compiled for.
#ifndef TDDA_LONG_WAVES
#define TDDA_LONG_WAVES 2
#endif

#ifndef TDDA_EPT_C64
#define TDDA_EPT_C64 12
#endif
#ifndef TDDA_EPT_C64_SMALL
#define TDDA_EPT_C64_SMALL 6
#endif
#ifndef TDDA_EPT_C128_SMALL
#define TDDA_EPT_C128_SMALL 6
#endif
#ifndef TDDA_EPT_SPLIT
#define TDDA_EPT_SPLIT 256
#endif
#ifndef TDDA_EPT_C128
#define TDDA_EPT_C128 6
#endif



enum SolveStatus : int {
    ST_RUNNING = 0,
    ST_CONVERGED = 1,
    ST_BREAKDOWN = 2,
    ST_MAXIT = 3,
};

// Per right-hand-side BiCGSTAB scalars (always FP64, also for the complex64
// inner solver) -- the scalars of DDAModel::bicgstab, DDAModel.cpp:145-160.
//
// One iteration runs in two operator applications with their finishing
// launches: P1 forms p (and first applies the previous iteration's pending
// updates of x and r), P5 forms Ap with r0.Ap -> alpha; P1 forms t, P5 forms
// At with the dot products from which the finishing launch takes zeta, the
// new rho = r0.r and |r| of r = t - zeta At without r itself.  `pend` marks
// x += alpha p + zeta t, r = t - zeta At as pending: the next P1 applies it,
// or k_settle once the job has stopped.
struct SolverState {
    double alpha[2], zeta[2], rho[2], rho_prev[2];
    double enorm, rnorm, err, tol;
    int iter, maxit, status, first, pend;
    int prec;  // right Jacobi preconditioning: the solve runs on A D^-1, D = diag(al)
    int idle;  // not part of this solve phase (a refinement cycle after it converged)
};

// Partial sums per tile of a reducing pass (the At pass needs eight).
constexpr int kPart = 8;

template <typename T>
struct JobDesc {
    int Nx, Ny, Nz, Mx, My, Mz, Hx, Hy, Hz;
    int N;       // sites
    int pz;      // pencils per tile in P5
    int pz1;     // pencils per tile in P1
    int tfy;     // flattened (kz, c) columns per tile in P2/P4
    int tfx;     // flattened (kz, c) columns per tile in P3 (multiple of 3)
    int nfy;     // tiles per x-slab in P2/P4: ceil(3 Mz / tfy)
    int nfx;     // tiles per ky row in P3: ceil(3 Mz / tfx)
    int ps_z, ps_y, ps_x;  // LDS pencil strides
    int xreg;              // P3 runs the register-resident kernel (k_pxr)
    int yreg;              // P2 / P4 run the register-resident kernel (k_pyr)
    int xwide;             // k_pxr with 512 threads per workgroup
    int ywide;             // k_pyr with 512 threads per workgroup
    int fzy;               // P1 + P2 and P4 + P5 by the fused kernels k_pzy / k_pyz (one tile per x slab)
    int gidx, gsize;       // place in its group of right-hand sides of one operator (consecutive jobs
                           // in the table, equal tiles), whose x-pass tiles are interleaved; 0, 1 alone
    const int* cellmap;    // [Nx][Ny][Nz] -> site or -1; nullptr for the full box in bulk order
    const Cx<T>* green;
    const Cx<T>* twx;   // two-level twiddle tables (runtime-plan kernels)
    const Cx<T>* twy;
    const Cx<T>* twz;
    const Cx<T>* twfx;  // full tables of the L-th roots (compile-time kernels)
    const Cx<T>* twfy;
    const Cx<T>* twfz;
    FftPlan plx, ply, plz;
    Cx<T>* w1;
    Cx<T>* w2;
    // plain operator application
    const Cx<T>* in;
    Cx<T>* out;
    const Cx<T>* al;
    // solver vectors (3N each)
    Cx<T>* x;
    Cx<T>* r;
    Cx<T>* r0;
    Cx<T>* p;
    Cx<T>* ap;
    Cx<T>* t;
    Cx<T>* at;
    const Cx<T>* e;
    SolverState* st;
    double* part;   // [tiles][4] partial sums of the reducing passes
};

// P1 input modes
enum InMode : int { IN_PLAIN = 0, IN_P = 1, IN_T = 2, IN_X = 3 };
// P5 output modes
enum OutMode : int { OUT_PLAIN = 0, OUT_RES = 1, OUT_AP = 2, OUT_AT = 3 };

__device__ __forceinline__ int find_job(const int* __restrict__ tstart, int njobs, int b)
{
    int lo = 0, hi = njobs - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (tstart[mid] <= b) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Global memory through the global address space.  The vectors of a job
// are generic pointers in its descriptor, for which the compiler emits flat
// instructions; a flat load counts in both the vector-memory and the LDS
// wait counters, so each LDS wait of a transform would also wait for the
// prefetch of the next tile, and loads issued back to back get serialised.
#define TDDA_GLOBAL __attribute__((address_space(1)))
typedef float tdda_f2 __attribute__((ext_vector_type(2)));
typedef double tdda_d2 __attribute__((ext_vector_type(2)));
template <typename T>
struct Vec2;
template <>
struct Vec2<float> {
    using type = tdda_f2;
};
template <>
struct Vec2<double> {
    using type = tdda_d2;
};

template <typename T>
__device__ __forceinline__ Cx<T> gld(const Cx<T>* p)
{
    using V = typename Vec2<T>::type;
    const V v = *(const TDDA_GLOBAL V*)p;
    return Cx<T>{v.x, v.y};
}

template <typename T>
__device__ __forceinline__ void gst(Cx<T>* p, Cx<T> c)
{
    using V = typename Vec2<T>::type;
    V v;
    v.x = c.x;
    v.y = c.y;
    *(TDDA_GLOBAL V*)p = v;
}

__device__ __forceinline__ int gldi(const int* p) { return *(const TDDA_GLOBAL int*)p; }

// Wait for every outstanding vector-memory load (s_waitcnt vmcnt(0); the
// other counters untouched: expcnt 7, lgkmcnt 15).  Issued at the top of
// each persistent-loop iteration, where the prefetched registers are
// consumed on every path, so that the compiler knows no load is in flight
// when it reuses those registers for the next prefetch's addresses --
// otherwise it waits for all loads before each load of the prefetch.
__device__ __forceinline__ void prefetch_landed() { __builtin_amdgcn_s_waitcnt(0x0F70); }

__device__ __forceinline__ const TDDA_GLOBAL SolverState* gstate(const SolverState* st)
{
    return (const TDDA_GLOBAL SolverState*)st;
}

template <typename T>
__device__ __forceinline__ Cx<T> ld(const Cx<T>* p) { return *p; }

__device__ __forceinline__ double2 to_d(Cx<float> a) { return make_double2(a.x, a.y); }
__device__ __forceinline__ double2 to_d(Cx<double> a) { return make_double2(a.x, a.y); }

template <typename T>
__device__ __forceinline__ Cx<T> from_d(double re, double im) { return Cx<T>{T(re), T(im)}; }

__device__ __forceinline__ void cdiv(double ar, double ai, double br, double bi, double& cr, double& ci)
{
    // complex division with scaling (Smith's method)
    if (fabs(br) >= fabs(bi)) {
        if (br == 0.0 && bi == 0.0) {
            cr = ar / br;
            ci = ai / br;
            return;
        }
        const double rr = bi / br, den = br + bi * rr;
        cr = (ar + ai * rr) / den;
        ci = (ai - ar * rr) / den;
    } else {
        const double rr = br / bi, den = br * rr + bi;
        cr = (ar * rr + ai) / den;
        ci = (ai * rr - ar) / den;
    }
}

// Block-wide deterministic sum of NV doubles held per thread; result valid in
// thread 0.  Uses wavefront shuffles then LDS.
template <int NV>
__device__ __forceinline__ void block_sum(double* v, double* red)
{
#pragma unroll
    for (int k = 0; k < NV; k++) {
        double s = v[k];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
        v[k] = s;
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < NV; k++) red[wid * NV + k] = v[k];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const int nw = blockDim.x >> 6;
#pragma unroll
        for (int k = 0; k < NV; k++) {
            double s = 0.0;
            for (int w = 0; w < nw; w++) s += red[w * NV + k];
            v[k] = s;
        }
    }
}

// Stores this tile's kPart partial sums (valid in thread 0 after block_sum).
// The totals are formed by k_finish in a separate launch: a last-workgroup
// ticket inside the pass would need a device-scope release fence per tile,
// which on a multi-XCD GPU writes back the XCD's L2 each time.
__device__ __forceinline__ void store_partials(double* part, int tile, const double* v)
{
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < kPart; k++) *(TDDA_GLOBAL double*)(part + tile * kPart + k) = v[k];
    }
}

__device__ __forceinline__ bool job_inactive(const SolverState* st)
{
    return st != nullptr && gstate(st)->status != ST_RUNNING;
}

// a / b
template <typename T>
__device__ __forceinline__ Cx<T> cdivv(Cx<T> a, Cx<T> b)
{
    const T inv = T(1) / (b.x * b.x + b.y * b.y);
    return Cx<T>{(a.x * b.x + a.y * b.y) * inv, (a.y * b.x - a.x * b.y) * inv};
}

// Dot products of the At pass: conj(At) t, conj(At) At (zeta,
// DDAModel.cpp:225), and |t|^2, conj(r0) t, conj(r0) At, from which the
// finishing launch forms r0.r and |r|^2 of r = t - zeta At (DDAModel.cpp:
// 230-232) without a pass over r.
template <typename T>
__device__ __forceinline__ void at_sums(double* acc, Cx<T> at, Cx<T> tv, Cx<T> r0)
{
    const double2 a = to_d(at), b = to_d(tv), c = to_d(r0);
    acc[0] += a.x * b.x + a.y * b.y;  // conj(At) t
    acc[1] += a.x * b.y - a.y * b.x;
    acc[2] += a.x * a.x + a.y * a.y;  // conj(At) At
    acc[3] += b.x * b.x + b.y * b.y;  // |t|^2
    acc[4] += c.x * b.x + c.y * b.y;  // conj(r0) t
    acc[5] += c.x * b.y - c.y * b.x;
    acc[6] += c.x * a.x + c.y * a.y;  // conj(r0) At
    acc[7] += c.x * a.y - c.y * a.x;
}

// ---------------------------------------------------------------------------
// Operator passes.  Every pass is a persistent kernel: a workgroup walks the
// flattened tile list (all jobs) with stride gridDim.x and software-pipelines
// it -- iteration i moves the prefetched inputs of tile i-1 into LDS, issues
// the global loads of tile i into registers, then transforms tile i-1, so the
// loads overlap the transform.  EP = ept_len<T>(L) is the number of
// prefetched elements per thread; tiles hold at most kThreads * EP elements
// (P1: a third of that per input vector).
// ---------------------------------------------------------------------------
template <typename T>
struct Ept;
template <>
struct Ept<double> {
    static constexpr int big = TDDA_EPT_C128;
    static constexpr int small = TDDA_EPT_C128_SMALL;
};
template <>
struct Ept<float> {
    static constexpr int big = TDDA_EPT_C64;
    static constexpr int small = TDDA_EPT_C64_SMALL;
};

// Prefetched elements per thread of a pass whose FFT length is L (0: the
// runtime-plan kernel, which takes the larger setting).  Pencils shorter
// than TDDA_EPT_SPLIT points use smaller tiles (less LDS per workgroup, more
// resident workgroups); longer pencils keep larger tiles so that the x pass
// still reads rows of several consecutive columns.
template <typename T>
__host__ __device__ constexpr int ept_len(int L)
{
    return (L == 0 || L >= TDDA_EPT_SPLIT) ? Ept<T>::big : Ept<T>::small;
}

// Prefetch registers: real and imaginary parts in separate scalar arrays
// (an array of 16-byte complex structs is left in scratch memory by the
// compiler; scalar arrays are promoted to registers).
template <typename T, int N>
struct Pre {
    T re[N], im[N];
    __device__ __forceinline__ Cx<T> get(int k) const { return Cx<T>{re[k], im[k]}; }
    __device__ __forceinline__ void set(int k, Cx<T> v)
    {
        re[k] = v.x;
        im[k] = v.y;
    }
};

// Walk of one thread over the elements e = threadIdx.x + k kThreads of a
// tile viewed as rows of n columns (row = e / n, col = e % n): one division
// per tile, then an add with carry per step instead of a division per
// element.
struct RowWalk {
    int row, col, drow, dcol, n;
    __device__ __forceinline__ RowWalk(int n_, float rn) : n(n_)
    {
        fdivmod((int)threadIdx.x, n_, rn, row, col);
        fdivmod(kThreads, n_, rn, drow, dcol);
    }
    __device__ __forceinline__ void next()
    {
        col += dcol;
        row += drow;
        const bool wrap = col >= n;
        col = wrap ? col - n : col;
        row = wrap ? row + 1 : row;
    }
};

// (z, c) of a flattened (z, component) column index
__device__ __forceinline__ void split3(int zc, int& z, int& c)
{
    z = (int)((unsigned)zc / 3u);
    c = zc - 3 * z;
}

// Largest pencil count of a tile of length L (compile-time lengths).
template <typename T>
__host__ __device__ constexpr int np_max(int L)
{
    return L > 0 ? (kThreads * ept_len<T>(L)) / L : 0;
}

// Whether an in-place transform of np pencils of length L keeps few enough
// values in registers: up to 8 complex values per lane; complex64
// power-of-two lengths up to 16 (other lengths that need more spill
// registers to scratch).
template <typename T>
__host__ __device__ constexpr bool ip_fits(int L, int np)
{
    const int v = ct_ip_values(L, np);
    return v <= 8 || (sizeof(T) == 4 && v <= 16 && (L & (L - 1)) == 0);
}

// x pass (P3) tile geometry.  For pencils of 120 points or more, where the
// in-place transform's registers allow it, a tile holds 12 pencils (three
// per wavefront), else 6 (a multiple of 3 for the Green product), in one
// LDS buffer, and the prefetch holds only the Nx <= (L + 1) / 2 loaded rows.
// Otherwise (and for shorter pencils, whose butterflies would fill too few
// lanes) the tile is sized like the other passes (np_max pencils, two LDS
// buffers).
template <typename T>
__host__ __device__ constexpr int px_inplace_pencils(int L)
{
    if (L < 120) return 0;
    if (ip_fits<T>(L, 12)) return 12;
    if (ip_fits<T>(L, 6)) return 6;
    return 0;
}
template <typename T>
__host__ __device__ constexpr bool px_inplace(int L)
{
    return px_inplace_pencils<T>(L) > 0;
}
// Waves per SIMD the x-pass kernel is compiled for: the in-place complex128
// kernels get the registers of three (with those of four they spill), all
// others those of TDDA_PASS_WAVES.
template <typename T>
__host__ __device__ constexpr int px_waves(int L)
{
    return (px_inplace<T>(L) && sizeof(T) == 8) ? 3 : TDDA_PASS_WAVES;
}
template <typename T>
__host__ __device__ constexpr int px_np(int L)
{
    return px_inplace<T>(L) ? px_inplace_pencils<T>(L) : np_max<T>(L);
}
template <typename T>
__host__ __device__ constexpr int px_ept(int L)
{
    return px_inplace<T>(L) ? (((L + 1) / 2) * px_inplace_pencils<T>(L) + kThreads - 1) / kThreads : ept_len<T>(L);
}

// z passes (P1/P5): tiles keep their size (np_max pencils) and are
// transformed in place in one LDS buffer where the registers allow it.
template <typename T>
__host__ __device__ constexpr bool pz_inplace(int L)
{
    return L > 0 && np_max<T>(L) >= 3 && ip_fits<T>(L, np_max<T>(L));
}

// y pass (P2/P4) tile geometry, in the same way: where registers allow, a
// tile holds kPyInplacePencils[precision] pencils (128-byte rows of the
// strided intermediate) in one LDS buffer, transformed in place, and the
// prefetch holds only the rows the pass reads (Ny <= (L + 1) / 2 forward, all
// L inverse).
template <typename T>
__host__ __device__ constexpr int py_inplace_pencils()
{
    return sizeof(T) == 8 ? 8 : 16;
}
template <typename T>
__host__ __device__ constexpr bool py_inplace(int L)
{
    return L >= 120 && ip_fits<T>(L, py_inplace_pencils<T>());
}
template <typename T>
__host__ __device__ constexpr int py_np(int L)
{
    return py_inplace<T>(L) ? py_inplace_pencils<T>() : np_max<T>(L);
}
template <typename T, bool INV>
__host__ __device__ constexpr int py_ept(int L)
{
    return py_inplace<T>(L) ? ((INV ? L : (L + 1) / 2) * py_inplace_pencils<T>() + kThreads - 1) / kThreads
                            : ept_len<T>(L);
}

// LDS slots of the largest data tile the host plans for a compile-time pass
// kernel of length L (kind 0: z, 1: y, 2: x), as plan_tiles in
// tdda_capi.hip does with an unclipped third axis.
template <typename T>
__host__ __device__ constexpr int max_tile_slots(int kind, int L)
{
    const int cap = kThreads * ept_len<T>(L);
    const int ps = L + (L >> pad_shift<T>()) + 1;
    const int unit = 128 / (int)sizeof(Cx<T>);
    if (kind == 0) {
        const int pz = cap / (3 * L) < 64 ? cap / (3 * L) : 64;
        return (pz_inplace<T>(L) ? 1 : 2) * 3 * pz * ps;
    }
    if (kind == 1) {
        if (py_inplace<T>(L)) return py_inplace_pencils<T>() * ps;
        int ty = cap / L;
        if (ty >= unit) ty = (ty / unit) * unit;
        return 2 * ty * ps;
    }
    if (px_inplace<T>(L)) return px_inplace_pencils<T>(L) * ps;
    int tx = cap / L < 3 * (8 * kThreads / L) ? cap / L : 3 * (8 * kThreads / L);
    tx = (tx / 3) * 3;
    if (tx >= 3 * unit) tx = (tx / (3 * unit)) * (3 * unit);
    return 2 * (tx > 3 ? tx : 3) * ps;
}

// LDS of one CU on gfx950, and the granule in which a workgroup's LDS is
// allocated.
__host__ __device__ constexpr int lds_per_cu() { return 160 * 1024; }
__host__ __device__ constexpr int lds_granule() { return 512; }
// Workgroups that fit in one CU's LDS when each takes `bytes`.
__host__ __device__ constexpr int lds_resident_workgroups(size_t bytes)
{
    const size_t g = ((bytes + lds_granule() - 1) / lds_granule()) * lds_granule();
    return g == 0 ? 1 << 20 : (int)(lds_per_cu() / g);
}

// Whether a compile-time pass kernel keeps the full table of the L-th roots
// in LDS (direct twiddle reads, fewer instructions per butterfly) or the
// two-level table: the full table is used unless it would lower the
// workgroups resident per CU at the largest tile (512 B allowed for a
// kernel's static reduction buffer).
template <typename T>
__host__ __device__ constexpr bool pass_full_twiddles(int kind, int L)
{
    const int data = max_tile_slots<T>(kind, L) * (int)sizeof(Cx<T>) + 512;
    const int full = pass_tw_slots(true, L, ct_tb(L)) * (int)sizeof(Cx<T>);
    const int two = pass_tw_slots(false, L, ct_tb(L)) * (int)sizeof(Cx<T>);
    return lds_resident_workgroups(data + full) >= lds_resident_workgroups(data + two);
}

// Transform of a pass tile: compile-time length LC > 0, or the runtime plan.
template <typename T, int LC, bool INV, int PRUNE = 0, bool FULLTW = false>
__device__ __forceinline__ Cx<T>* pass_fft(Cx<T>* a, Cx<T>* b, int np, int ps, const FftPlan& pl, const Cx<T>* tw)
{
    if constexpr (LC > 0) {
        return ct_fft<T, LC, INV, np_max<T>(LC), PRUNE, FULLTW>(a, b, np, tw);
    } else {
        return pp_fft<T, INV>(a, b, np, ps, pl, tw);
    }
}

// Upper end of the zero padding a forward pass must write in LDS: the
// pruned first stage of the compile-time transform reads only below
// ct_zin_limit(L); the runtime-plan transform reads everything.
template <int LC>
__device__ __forceinline__ int zero_fill_end(int M)
{
    if constexpr (LC > 0) return min(M, ct_zin_limit(LC));
    else return M;
}

// Geometry of a z-pass tile (P1 / P5): pencils [q0, q0 + np) of job `job`.
struct ZTile {
    int job, jt, q0, np;
    bool live;
};

template <typename T>
__device__ __forceinline__ ZTile z_tile(const JobDesc<T>* __restrict__ jobs, const int* __restrict__ tstart,
                                        int njobs, int tl, int pz_sel, bool check_state)
{
    ZTile z;
    z.job = find_job(tstart, njobs, tl);
    const JobDesc<T>* J = jobs + z.job;
    const int pz = pz_sel ? J->pz1 : J->pz;
    z.jt = tl - tstart[z.job];
    z.q0 = z.jt * pz;
    z.np = min(pz, J->Nx * J->Ny - z.q0);
    z.live = !(check_state && job_inactive(J->st));
    return z;
}

// Geometry of a strided-pass tile (P2/P4: outer = x; P3: outer = ky).
struct STile {
    int job, outer, f0, nf;
    bool live;
};

template <typename T>
__device__ __forceinline__ STile s_tile(const JobDesc<T>* __restrict__ jobs, const int* __restrict__ tstart,
                                        int njobs, int tl, bool xpass, bool check_state)
{
    STile t;
    t.job = find_job(tstart, njobs, tl);
    const JobDesc<T>* J = jobs + t.job;
    const int nfb = xpass ? J->nfx : J->nfy;
    const int tf = xpass ? J->tfx : J->tfy;
    int o, fb;
    if (xpass) {
        // the x-pass tiles of a group of right-hand sides of one operator
        // (consecutive jobs with equal tile counts) are interleaved
        const int first = t.job - J->gidx;
        int k;
        xtile_map(tl - tstart[first], J->My, nfb, J->gsize, o, fb, k);
        t.job = first + k;
        J = jobs + t.job;
    } else {
        ytile_map(tl - tstart[t.job], J->Nx, nfb, o, fb);
    }
    t.f0 = fb * tf;
    t.outer = o;
    t.nf = min(tf, J->Mz * 3 - t.f0);
    t.live = !(check_state && job_inactive(J->st));
    return t;
}

// The x-pass tile tl from a tile map the host built with s_tile's
// arithmetic: (job, ky, f0, nf) per tile, one scalar load instead of the
// job search and the interleaving divisions.
template <typename T>
__device__ __forceinline__ STile m_tile(const JobDesc<T>* __restrict__ jobs, const int4* __restrict__ tmap, int tl,
                                        bool check_state)
{
    const int4 m = tmap[tl];
    STile t;
    t.job = m.x;
    t.outer = m.y;
    t.f0 = m.z;
    t.nf = m.w;
    t.live = !(check_state && job_inactive(jobs[t.job].st));
    return t;
}

// P1: z-forward.  Tile = pz1 consecutive (x, y) pencils.
template <typename T, int MODE, int LC>
__global__ void TDDA_PASS_BOUNDS k_p1(const JobDesc<T>* __restrict__ jobs, const int* __restrict__ tstart,
                                      int njobs)
{
    extern __shared__ __align__(16) char smem[];
    constexpr bool FT = LC > 0 && pass_full_twiddles<T>(0, LC);
    constexpr bool IP = LC > 0 && pz_inplace<T>(LC);  // one LDS buffer, in-place transforms
    // IN_P: t, At, x, p, Ap (the pending update, then p); the first step: r.
    // IN_P / IN_T with right Jacobi preconditioning also al (the operator
    // input is p / al or t / al), in the last slot.
    constexpr int NV = (MODE == IN_P) ? 6 : ((MODE == IN_T) ? 3 : 1);
    constexpr int EP = ept_len<T>(LC) / 3;
    const int ntot = tstart[njobs];
    Pre<T, EP> pr[NV];
    ZTile cur;
    cur.live = false;
    for (int tl = blockIdx.x;; tl += gridDim.x) {
        prefetch_landed();
        // 1. inputs of the current tile -> LDS (fused BiCGSTAB vector updates)
        if (cur.live) {
            const JobDesc<T>* J = jobs + cur.job;
            const int Nz = J->Nz, Mz = J->Mz, ps = J->ps_z;
            Cx<T>* twl = reinterpret_cast<Cx<T>*>(smem);
            Cx<T>* A = twl + pass_tw_slots(FT, Mz, J->plz.tb);
            load_pass_twiddles<T, FT>(twl, J->twz, J->twfz, J->plz);
            Cx<T> sa{T(0), T(0)}, sb{T(0), T(0)}, sc{T(0), T(0)};
            bool first = false;
            const bool prec = (MODE == IN_P || MODE == IN_T) && gstate(J->st)->prec != 0;
            if (MODE == IN_P) {
                const TDDA_GLOBAL SolverState* st = gstate(J->st);
                first = st->first != 0;
                if (!first) {
                    // beta = (alpha / zeta) * rho / rho_prev   (DDAModel.cpp:218)
                    double q1r, q1i, q2r, q2i;
                    cdiv(st->alpha[0], st->alpha[1], st->zeta[0], st->zeta[1], q1r, q1i);
                    cdiv(st->rho[0], st->rho[1], st->rho_prev[0], st->rho_prev[1], q2r, q2i);
                    sa = from_d<T>(q1r * q2r - q1i * q2i, q1r * q2i + q1i * q2r);
                    sb = from_d<T>(st->zeta[0], st->zeta[1]);
                    sc = from_d<T>(st->alpha[0], st->alpha[1]);
                }
            } else if (MODE == IN_T) {
                sa = from_d<T>(gstate(J->st)->alpha[0], gstate(J->st)->alpha[1]);
            }
            const int nin = cur.np * Nz * 3;
            const int base = cur.q0 * Nz;
            RowWalk w(Nz * 3, 1.0f / (float)(Nz * 3));
#pragma unroll
            for (int k = 0; k < EP; k++, w.next()) {
                const int e = threadIdx.x + k * kThreads;
                if (e >= nin) continue;
                const int pl = w.row;
                int z, c;
                split3(w.col, z, c);
                Cx<T> v = pr[0].get(k);
                if (MODE == IN_P || MODE == IN_T) {
                    int idx = -1;
                    if (J->cellmap) {
                        const int site = gldi(J->cellmap + base + pl * Nz + z);
                        if (site >= 0) idx = site * 3 + c;
                    } else {
                        idx = base * 3 + e;
                    }
                    if (MODE == IN_P) {
                        if (!first) {
                            // the previous iteration's P += alpha p + z (z = zeta r - alpha u =
                            // zeta t) and r = t - zeta At (DDAModel.cpp:226-230), then
                            // p = r + beta (p - u), u = zeta Ap   (DDAModel.cpp:219)
                            const Cx<T> tv = pr[0].get(k), pv = pr[3].get(k);
                            const Cx<T> rn = csub(tv, cmul(sb, pr[1].get(k)));
                            v = cadd(rn, cmul(sa, csub(pv, cmul(sb, pr[4].get(k)))));
                            if (idx >= 0) {
                                gst(J->r + idx, rn);
                                gst(J->x + idx, cadd(pr[2].get(k), cadd(cmul(sc, pv), cmul(sb, tv))));
                            }
                        }
                        if (idx >= 0) gst(J->p + idx, v);
                    } else {
                        // t = r - alpha Ap   (DDAModel.cpp:222)
                        v = csub(pr[0].get(k), cmul(sa, pr[1].get(k)));
                        if (idx >= 0) gst(J->t + idx, v);
                    }
                    if (prec && idx >= 0) v = cdivv(v, pr[NV - 1].get(k));
                }
                A[(pl * 3 + c) * ps + lpad<T>(z)] = v;
            }
            const int npad = zero_fill_end<LC>(Mz) - Nz;
            if (npad > 0) {
                for (RowWalk w(npad, 1.0f / (float)npad); w.row < cur.np * 3; w.next())
                    A[w.row * ps + lpad<T>(Nz + w.col)] = Cx<T>{T(0), T(0)};
            }
        }
        // 2. prefetch of the next tile
        ZTile nxt;
        nxt.live = false;
        if (tl < ntot) {
            nxt = z_tile(jobs, tstart, njobs, tl, 1, MODE != IN_PLAIN);
            if (nxt.live) {
                const JobDesc<T>* J = jobs + nxt.job;
                const bool pfirst = MODE == IN_P && gstate(J->st)->first != 0;
                const bool pprec = (MODE == IN_P || MODE == IN_T) && gstate(J->st)->prec != 0;
                const int Nz = J->Nz;
                const int nin = nxt.np * Nz * 3;
                const int base = nxt.q0 * Nz;
                const float rNz3 = 1.0f / (float)(Nz * 3);
#pragma unroll
                for (int k = 0; k < EP; k++) {
                    const int e = threadIdx.x + k * kThreads;
                    if (e >= nin) continue;
                    int idx;
                    if (J->cellmap) {
                        int pl, rem, z, c;
                        fdivmod(e, Nz * 3, rNz3, pl, rem);
                        split3(rem, z, c);
                        const int site = gldi(J->cellmap + base + pl * Nz + z);
                        idx = site >= 0 ? site * 3 + c : -1;
                    } else {
                        idx = base * 3 + e;  // bulk order: the tile is one contiguous run
                    }
                    if (idx < 0) {
#pragma unroll
                        for (int v = 0; v < NV; v++) pr[v].set(k, Cx<T>{T(0), T(0)});
                        continue;
                    }
                    if (MODE == IN_PLAIN) pr[0].set(k, gld(J->in + idx));
                    else if (MODE == IN_X) pr[0].set(k, gld(J->x + idx));
                    else if (MODE == IN_P) {
                        if (pfirst) {
                            pr[0].set(k, gld(J->r + idx));
                        } else {
                            pr[0].set(k, gld(J->t + idx));
                            pr[1].set(k, gld(J->at + idx));
                            pr[2].set(k, gld(J->x + idx));
                            pr[3].set(k, gld(J->p + idx));
                            pr[4].set(k, gld(J->ap + idx));
                        }
                    } else {
                        pr[0].set(k, gld(J->r + idx));
                        pr[1].set(k, gld(J->ap + idx));
                    }
                    if ((MODE == IN_P || MODE == IN_T) && pprec) pr[NV - 1].set(k, gld(J->al + idx));
                }
            }
        }
        // 3. transform of the current tile
        if (cur.live) {
            const JobDesc<T>* J = jobs + cur.job;
            const int Mz = J->Mz, ps = J->ps_z;
            Cx<T>* twl = reinterpret_cast<Cx<T>*>(smem);
            Cx<T>* A = twl + pass_tw_slots(FT, Mz, J->plz.tb);
            Cx<T>* B = IP ? A : A + 3 * J->pz1 * ps;
            __syncthreads();
            const Cx<T>* res;
            if constexpr (IP) res = ct_fft_ip<T, LC, false, np_max<T>(LC), kPruneIn, FT>(A, cur.np * 3, twl);
            else res = pass_fft<T, LC, false, kPruneIn, FT>(A, B, cur.np * 3, ps, J->plz, twl);
            const int nout = cur.np * Mz * 3;
            Cx<T>* w1 = J->w1 + (size_t)cur.q0 * Mz * 3;
            RowWalk w(Mz * 3, 1.0f / (float)(Mz * 3));
            for (int e = threadIdx.x; e < nout; e += kThreads, w.next()) {
                int kz, c;
                split3(w.col, kz, c);
                gst(w1 + e, res[(w.row * 3 + c) * ps + lpad<T>(kz)]);
            }
            __syncthreads();
        }
        if (tl >= ntot) break;
        cur = nxt;
    }
}

// P2 / P4: y pass.  Tile = (x, tfy consecutive flattened (kz, c) columns).
// FWD: w1[x][y<Ny][.] -> w2[x][ky<My][.];  INV: w2[x][ky][.] -> w1[x][y<Ny][.]
template <typename T, bool INV, int LC>
__global__ void TDDA_PASS_BOUNDS k_py(const JobDesc<T>* __restrict__ jobs, const int* __restrict__ tstart,
                                      int njobs, int solver)
{
    extern __shared__ __align__(16) char smem[];
    constexpr bool FT = LC > 0 && pass_full_twiddles<T>(1, LC);
    constexpr bool IP = LC > 0 && py_inplace<T>(LC);  // one LDS buffer, in-place transforms
    constexpr int EP = py_ept<T, INV>(LC);
    const int ntot = tstart[njobs];
    Pre<T, EP> pr;
    STile cur;
    cur.live = false;
    for (int tl = blockIdx.x;; tl += gridDim.x) {
        prefetch_landed();
        if (cur.live) {
            const JobDesc<T>* J = jobs + cur.job;
            const int Ny = J->Ny, My = J->My, ps = J->ps_y, nf = cur.nf;
            Cx<T>* twl = reinterpret_cast<Cx<T>*>(smem);
            Cx<T>* A = twl + pass_tw_slots(FT, My, J->ply.tb);
            load_pass_twiddles<T, FT>(twl, J->twy, J->twfy, J->ply);
            const int nin = INV ? My : Ny;
            const float rnf = 1.0f / (float)nf;
            RowWalk w(nf, rnf);
#pragma unroll
            for (int k = 0; k < EP; k++, w.next())
                if (w.row < nin) A[w.col * ps + lpad<T>(w.row)] = pr.get(k);
            if (!INV) {
                for (RowWalk z(nf, rnf); z.row < zero_fill_end<LC>(My) - Ny; z.next())
                    A[z.col * ps + lpad<T>(Ny + z.row)] = Cx<T>{T(0), T(0)};
            }
        }
        STile nxt;
        nxt.live = false;
        if (tl < ntot) {
            nxt = s_tile(jobs, tstart, njobs, tl, false, solver != 0);
            if (nxt.live) {
                const JobDesc<T>* J = jobs + nxt.job;
                const int F = J->Mz * 3, nf = nxt.nf;
                const int nin = INV ? J->My : J->Ny;
                const Cx<T>* src =
                    (INV ? J->w2 + (size_t)nxt.outer * J->My * F : J->w1 + (size_t)nxt.outer * J->Ny * F) + nxt.f0;
                RowWalk w(nf, 1.0f / (float)nf);
#pragma unroll
                for (int k = 0; k < EP; k++, w.next())
                    if (w.row < nin) pr.set(k, gld(src + (w.row * F + w.col)));
            }
        }
        if (cur.live) {
            const JobDesc<T>* J = jobs + cur.job;
            const int Ny = J->Ny, My = J->My, ps = J->ps_y, nf = cur.nf, F = J->Mz * 3;
            Cx<T>* twl = reinterpret_cast<Cx<T>*>(smem);
            Cx<T>* A = twl + pass_tw_slots(FT, My, J->ply.tb);
            Cx<T>* B = IP ? A : A + J->tfy * ps;
            __syncthreads();
            const Cx<T>* res;
            if constexpr (IP) res = ct_fft_ip<T, LC, INV, py_inplace_pencils<T>(), INV ? kPruneOut : kPruneIn, FT>(A, nf, twl);
            else res = pass_fft<T, LC, INV, INV ? kPruneOut : kPruneIn, FT>(A, B, nf, ps, J->ply, twl);
            const int nout = INV ? Ny : My;
            Cx<T>* dst =
                (INV ? J->w1 + (size_t)cur.outer * Ny * F : J->w2 + (size_t)cur.outer * My * F) + cur.f0;
            for (RowWalk w(nf, 1.0f / (float)nf); w.row < nout; w.next())
                gst(dst + (w.row * F + w.col), res[w.col * ps + lpad<T>(w.row)]);
            __syncthreads();
        }
        if (tl >= ntot) break;
        cur = nxt;
    }
}

// Register-resident y pass (P2 / P4) for lengths L = L1 * L2 with both
// factors in the small-DFT set (pxr_l1, L1 >= L2); same four-step split as
// the x pass (y = n1 + L1 n2, ky = k2 + L2 k1), with one LDS exchange:
//   forward: thread (col, n1) loads w1 rows y = n1 + L1 n2 < Ny, DFT over n2,
//            times w_L^(n1 k2), to LDS slot n1 L2 + k2; thread (col, k2)
//            reads its L1 slots, DFT over n1, stores w2 rows ky = k2 + L2 k1;
//   inverse: thread (col, k2) loads w2 rows k2 + L2 k1, inverse DFT over k1,
//            times conj(w_L^(n1 k2)), to LDS; thread (col, n1) reads its L2
//            slots, inverse DFT over k2, stores w1 rows y = n1 + L1 n2 < Ny.
// A tile is (x, nf consecutive (kz, c) columns), nf * L1 <= NT (threads of
// the workgroup: 256, or 512 for factors above 16, whose 256-thread tiles
// would read rows of 8 columns).
template <typename T>
__host__ __device__ constexpr int pyr_nf(int L, int nt = kThreads)
{
    return pxr_l1(L) <= 0 ? 0 : (nt / pxr_l1(L) >= 8 ? ((nt / pxr_l1(L)) / 8) * 8 : nt / pxr_l1(L));
}

// complex128 from 256 points: three waves, two where the inner factor of
// its split reaches 20 (400 = 20 x 20 .. 540 = 27 x 20 spill at three)
template <typename T>
__host__ __device__ constexpr int wide_c128_waves(int L)
{
    return (pxr_l1(L) > 0 && L / pxr_l1(L) >= 20) ? 2 : 3;
}

// complex128 from 256 points needs the registers of three waves (with four
// it spills), two where its inner factor reaches 20; all others fit four.
template <typename T>
__host__ __device__ constexpr int pyr_waves(int L)
{
    return (sizeof(T) == 8 && L >= 256) ? wide_c128_waves<T>(L) : pxr_l1(L) > 16 ? TDDA_LONG_WAVES : 4;
}

template <typename T, bool INV, int LC, int NT = kThreads>
__global__ void __launch_bounds__(NT, pyr_waves<T>(LC)) k_pyr(const JobDesc<T>* __restrict__ jobs,
                                                     const int* __restrict__ tstart, int njobs, int solver,
                                                     const int4* __restrict__ tmap)
{
    constexpr int L1 = pxr_l1(LC), L2 = LC / L1, H2 = (L2 + 1) / 2;
    static_assert(L1 > 0 && L1 >= L2, "length without a register split");
    extern __shared__ __align__(16) char smem[];
    Cx<T>* tw = reinterpret_cast<Cx<T>*>(smem);  // tw[q] = exp(-2 pi i q / L)
    Cx<T>* A = tw + pass_tw_slots(true, LC, 0);
    const int ntot = tstart[njobs];
    for (int i = threadIdx.x; i < LC; i += NT) tw[i] = gld(jobs[0].twfy + i);
    __syncthreads();
    const int task = threadIdx.x;
    for (int tl = blockIdx.x; tl < ntot; tl += gridDim.x) {
        const STile t = tmap ? m_tile(jobs, tmap, tl, solver != 0) : s_tile(jobs, tstart, njobs, tl, false, solver != 0);
        if (!t.live) continue;
        const JobDesc<T>* J = jobs + t.job;
        const int Ny = J->Ny, F = 3 * J->Mz, nf = t.nf, ps = J->ps_y;
        Cx<T>* w1 = J->w1 + (size_t)t.outer * Ny * F + t.f0;
        Cx<T>* w2 = J->w2 + (size_t)t.outer * LC * F + t.f0;
        const int q = task / nf, col = task - q * nf;  // (n1 or k2, col)
        if constexpr (!INV) {
            if (q < L1) {
                const int n1 = q;
                Cx<T> v[L2];
                sfor<0, L2>([&](auto c2) {
                    constexpr int n2 = decltype(c2)::value;
                    if constexpr (n2 >= H2) {
                        v[n2] = Cx<T>{T(0), T(0)};
                    } else {
                        const int y = n1 + L1 * n2;
                        v[n2] = y < Ny ? gld(w1 + (y * F + col)) : Cx<T>{T(0), T(0)};
                    }
                });
                sdft<T, L2, false>(v);
                Cx<T>* row = A + col * ps + n1 * L2;
                row[0] = v[0];
                sfor<1, L2>([&](auto c2) {
                    constexpr int k2 = decltype(c2)::value;
                    row[k2] = cmul(v[k2], tw[n1 * k2]);
                });
            }
            __syncthreads();
            if (q < L2) {
                const int k2 = q;
                Cx<T> v[L1];
                const Cx<T>* cp = A + col * ps + k2;
                sfor<0, L1>([&](auto c1) { v[decltype(c1)::value] = cp[decltype(c1)::value * L2]; });
                sdft<T, L1, false>(v);
                sfor<0, L1>([&](auto c1) {
                    constexpr int k1 = decltype(c1)::value;
                    gst(w2 + ((k2 + L2 * k1) * F + col), v[k1]);
                });
            }
        } else {
            if (q < L2) {
                const int k2 = q;
                Cx<T> v[L1];
                sfor<0, L1>([&](auto c1) {
                    constexpr int k1 = decltype(c1)::value;
                    v[k1] = gld(w2 + ((k2 + L2 * k1) * F + col));
                });
                sdft<T, L1, true>(v);
                Cx<T>* cp = A + col * ps + k2;
                cp[0] = v[0];
                sfor<1, L1>([&](auto c1) {
                    constexpr int n1 = decltype(c1)::value;
                    cp[n1 * L2] = cmulc(v[n1], tw[n1 * k2]);
                });
            }
            __syncthreads();
            if (q < L1) {
                const int n1 = q;
                Cx<T> v[L2];
                const Cx<T>* row = A + col * ps + n1 * L2;
                sfor<0, L2>([&](auto c2) { v[decltype(c2)::value] = row[decltype(c2)::value]; });
                sdft<T, L2, true>(v);
                sfor<0, H2>([&](auto c2) {
                    constexpr int n2 = decltype(c2)::value;
                    const int y = n1 + L1 * n2;
                    if (y < Ny) gst(w1 + (y * F + col), v[n2]);
                });
            }
        }
        __syncthreads();
    }
}

// y = G b at one k point of the three component rows v[0], v[ps], v[2 ps].
template <typename T>
__device__ __forceinline__ void green_apply(Cx<T>* v, int ps, Cx<T> gxx, Cx<T> gxy, Cx<T> gxz, Cx<T> gyy, Cx<T> gyz,
                                            Cx<T> gzz)
{
    const Cx<T> bx = v[0], by = v[ps], bz = v[2 * ps];
    v[0] = cadd(cadd(cmul(gxx, bx), cmul(gxy, by)), cmul(gxz, bz));
    v[ps] = cadd(cadd(cmul(gxy, bx), cmul(gyy, by)), cmul(gyz, bz));
    v[2 * ps] = cadd(cadd(cmul(gxz, bx), cmul(gyz, by)), cmul(gzz, bz));
}

// P3: x pass fused with the Green product (kernel.cu:74-95).
// Tile = (ky, tfx consecutive flattened (kz, c) columns, tfx % 3 == 0).
template <typename T, int LC>
__global__ void __launch_bounds__(kThreads, px_waves<T>(LC)) k_px(const JobDesc<T>* __restrict__ jobs, const int* __restrict__ tstart,
                                      int njobs, int solver)
{
    extern __shared__ __align__(16) char smem[];
    constexpr bool FT = LC > 0 && pass_full_twiddles<T>(2, LC);
    constexpr bool IP = LC > 0 && px_inplace<T>(LC);  // one LDS buffer, in-place transforms
    constexpr int EP = px_ept<T>(LC);
    const int ntot = tstart[njobs];
    Pre<T, EP> pr;
    STile cur;
    cur.live = false;
    for (int tl = blockIdx.x;; tl += gridDim.x) {
        prefetch_landed();
        if (cur.live) {
            const JobDesc<T>* J = jobs + cur.job;
            const int Nx = J->Nx, Mx = J->Mx, ps = J->ps_x, nf = cur.nf;
            Cx<T>* twl = reinterpret_cast<Cx<T>*>(smem);
            Cx<T>* A = twl + pass_tw_slots(FT, Mx, J->plx.tb);
            load_pass_twiddles<T, FT>(twl, J->twx, J->twfx, J->plx);
            const float rnf = 1.0f / (float)nf;
            RowWalk w(nf, rnf);
#pragma unroll
            for (int k = 0; k < EP; k++, w.next())
                if (w.row < Nx) A[w.col * ps + lpad<T>(w.row)] = pr.get(k);
            for (RowWalk z(nf, rnf); z.row < zero_fill_end<LC>(Mx) - Nx; z.next())
                A[z.col * ps + lpad<T>(Nx + z.row)] = Cx<T>{T(0), T(0)};
        }
        STile nxt;
        nxt.live = false;
        if (tl < ntot) {
            nxt = s_tile(jobs, tstart, njobs, tl, true, solver != 0);
            if (nxt.live) {
                const JobDesc<T>* J = jobs + nxt.job;
                const int F = J->Mz * 3, nf = nxt.nf;
                const int sx = J->My * F;  // one slab row: < 2^31 elements
                const Cx<T>* base = J->w2 + (size_t)nxt.outer * F + nxt.f0;
                const int Nx = J->Nx;
                RowWalk w(nf, 1.0f / (float)nf);
#pragma unroll
                for (int k = 0; k < EP; k++, w.next())
                    if (w.row < Nx) pr.set(k, gld(base + (w.row * sx + w.col)));
            }
        }
        if (cur.live) {
            const JobDesc<T>* J = jobs + cur.job;
            const int Nx = J->Nx, Mx = J->Mx, My = J->My, Mz = J->Mz, F = Mz * 3;
            const int ps = J->ps_x, nf = cur.nf, ky = cur.outer, f0 = cur.f0;
            Cx<T>* twl = reinterpret_cast<Cx<T>*>(smem);
            Cx<T>* A = twl + pass_tw_slots(FT, Mx, J->plx.tb);
            Cx<T>* B = IP ? A : A + J->tfx * ps;
            __syncthreads();
            Cx<T>* res;
            if constexpr (IP) res = ct_fft_ip<T, LC, false, px_np<T>(LC), kPruneIn, FT>(A, nf, twl);
            else res = pass_fft<T, LC, false, kPruneIn, FT>(A, B, nf, ps, J->plx, twl);
            Cx<T>* other = (res == A) ? B : A;
            // Green product.  Octant fold per axis: k' = min(k, M - k); odd
            // components flip sign when k was folded.
            const int kyf = (ky > My / 2) ? My - ky : ky;
            const T sy = (ky > My / 2) ? T(-1) : T(1);
            const int nkz = nf / 3;
            const int Hy = J->Hy, Hz = J->Hz;
            // One Green entry serves kx = kxf and its mirror Mx - kxf (which
            // flips the sign of the xy and xz components).
            for (RowWalk w(nkz, 1.0f / (float)nkz); w.row <= Mx / 2; w.next()) {
                const int kxf = w.row, kl = w.col;
                const int kz = f0 / 3 + kl;
                const int kzf = (kz > Mz / 2) ? Mz - kz : kz;
                const T szv = (kz > Mz / 2) ? T(-1) : T(1);
                const Cx<T>* gp = J->green + ((size_t)(kxf * Hy + kyf) * Hz + kzf) * 6;
                const Cx<T> gxx = gld(gp), gyy = gld(gp + 3), gzz = gld(gp + 5);
                const Cx<T> gxy = cscale(gld(gp + 1), sy);
                const Cx<T> gxz = cscale(gld(gp + 2), szv);
                const Cx<T> gyz = cscale(gld(gp + 4), sy * szv);
                green_apply(res + (kl * 3) * ps + lpad<T>(kxf), ps, gxx, gxy, gxz, gyy, gyz, gzz);
                if (kxf != 0 && 2 * kxf != Mx)
                    green_apply(res + (kl * 3) * ps + lpad<T>(Mx - kxf), ps, gxx, cneg(gxy), cneg(gxz), gyy, gyz,
                                gzz);
            }
            __syncthreads();
            const Cx<T>* res2;
            if constexpr (IP) res2 = ct_fft_ip<T, LC, true, px_np<T>(LC), kPruneOut, FT>(res, nf, twl);
            else res2 = pass_fft<T, LC, true, kPruneOut, FT>(res, other, nf, ps, J->plx, twl);
            const int sx = My * F;
            Cx<T>* base = J->w2 + (size_t)ky * F + f0;
            for (RowWalk w(nf, 1.0f / (float)nf); w.row < Nx; w.next())
                gst(base + (w.row * sx + w.col), res2[w.col * ps + lpad<T>(w.row)]);
            __syncthreads();
        }
        if (tl >= ntot) break;
        cur = nxt;
    }
}

// Register-resident x pass (P3) for lengths L = L1 * L2 with both factors
// in the small-DFT set (pxr_l1, L1 >= L2).  A tile is (ky, nf consecutive
// (kz, c) columns), nf a multiple of 3 with nf * L1 <= NT (threads of the
// workgroup: 256, or 512 for tiles twice as wide -- longer HBM rows and
// Green row segments at the same waves per SIMD).  With
// x = n1 + L1 n2 and kx = k2 + L2 k1 (four-step split):
//   A : thread (col, n1) loads x[n1 + L1 n2] from w2 (n2 < ceil(L2/2): the
//       rest is zero padding), DFT over n2 in registers, times w_L^(n1 k2),
//       to LDS slot n1 L2 + k2;
//   B : thread (col, k2) reads the L1 slots of its k2, DFT over n1 in
//       registers, back to the same slots, now kx = k2 + L2 k1 (natural
//       order);
//   G : the symmetric Green product in place in LDS, one thread per
//       (kz, |kx|) pair of the tile (one octant load serves kx and L - kx);
//   B': thread (col, k2) reads its L1 slots, inverse DFT over k1, times
//       conj(w_L^(n1 k2)), back to the same slots;
//   A': thread (col, n1) reads its L2 slots, inverse DFT over k2, and stores
//       the x < Nx outputs to w2.
// Each element crosses LDS four times (the LDS-staged transforms cross it at
// every radix stage).  No prefetch: the resident workgroups of a CU overlap
// each other's loads.
template <typename T>
__host__ __device__ constexpr int pxr_nf(int L, int nt = kThreads)
{
    return pxr_l1(L) > 0 ? ((nt / pxr_l1(L)) / 3) * 3 : 0;
}

// Waves per SIMD k_pxr is compiled for, the most at which the compiler
// keeps it out of scratch: four (complex64 50-64 VGPRs, complex128 73-122),
// three for complex128 from 256 points (131-165 VGPRs; with four they
// spill), two where complex128's inner factor reaches 20 (400 = 20 x 20 ..
// 540 = 27 x 20 spill at three).
template <typename T>
__host__ __device__ constexpr int pxr_waves(int L)
{
    return (sizeof(T) == 8 && L >= 256) ? wide_c128_waves<T>(L) : pxr_l1(L) > 16 ? TDDA_LONG_WAVES : 4;
}

template <typename T, int LC, int NT>
__global__ void __launch_bounds__(NT, pxr_waves<T>(LC)) k_pxr(const JobDesc<T>* __restrict__ jobs,
                                                     const int* __restrict__ tstart, int njobs, int solver,
                                                     const int4* __restrict__ tmap)
{
    constexpr int L1 = pxr_l1(LC), L2 = LC / L1, H2 = (L2 + 1) / 2;
    static_assert(L1 > 0 && L1 >= L2, "length without a register split");
    extern __shared__ __align__(16) char smem[];
    Cx<T>* tw = reinterpret_cast<Cx<T>*>(smem);  // tw[q] = exp(-2 pi i q / L)
    Cx<T>* A = tw + pass_tw_slots(true, LC, 0);
    const int ntot = tstart[njobs];
    // the full table of the L-th roots depends on L only: one per launch
    for (int i = threadIdx.x; i < LC; i += NT) tw[i] = gld(jobs[0].twfx + i);
    __syncthreads();
    const int task = threadIdx.x;
    for (int tl = blockIdx.x; tl < ntot; tl += gridDim.x) {
        const STile t = tmap ? m_tile(jobs, tmap, tl, solver != 0) : s_tile(jobs, tstart, njobs, tl, true, solver != 0);
        if (!t.live) continue;
        const JobDesc<T>* J = jobs + t.job;
        const int Nx = J->Nx, My = J->My, Mz = J->Mz, F = 3 * Mz, nf = t.nf, ky = t.outer, ps = J->ps_x;
        const int sx = My * F;  // one slab row: < 2^31 elements
        Cx<T>* base = J->w2 + (size_t)ky * F + t.f0;
        const int qa = task / nf, ca = task - qa * nf;  // (n1, col) of steps A / A', (k2, col) of B / G
        // A: forward DFT over n2
        if (qa < L1) {
            const int n1 = qa;
            Cx<T> v[L2];
            sfor<0, L2>([&](auto c2) {
                constexpr int n2 = decltype(c2)::value;
                if constexpr (n2 >= H2) {
                    v[n2] = Cx<T>{T(0), T(0)};
                } else {
                    const int x = n1 + L1 * n2;
                    v[n2] = x < Nx ? gld(base + (x * sx + ca)) : Cx<T>{T(0), T(0)};
                }
            });
            sdft<T, L2, false>(v);
            Cx<T>* row = A + ca * ps + n1 * L2;
            row[0] = v[0];
            sfor<1, L2>([&](auto c2) {
                constexpr int k2 = decltype(c2)::value;
                row[k2] = cmul(v[k2], tw[n1 * k2]);
            });
        }
        __syncthreads();
        const bool actb = qa < L2;
        const int k2 = qa;
        Cx<T> v[L1];
        // B: forward DFT over n1; slot k2 + L2 k1 then holds kx = k2 + L2 k1
        if (actb) {
            Cx<T>* cp = A + ca * ps + k2;
            sfor<0, L1>([&](auto c1) { v[decltype(c1)::value] = cp[decltype(c1)::value * L2]; });
            sdft<T, L1, false>(v);
            sfor<0, L1>([&](auto c1) { cp[decltype(c1)::value * L2] = v[decltype(c1)::value]; });
        }
        __syncthreads();
        // G: the Green product in place in LDS, one thread per (kz, |kx|):
        // the six octant components serve kx and its mirror L - kx (which
        // flips the sign of xy and xz), the fold of ky and kz flips the
        // components odd in y or z
        {
            const int nkz = nf / 3, nw = nkz * (LC / 2 + 1);
            const int kyf = (ky > My / 2) ? My - ky : ky;
            const T sy = (ky > My / 2) ? T(-1) : T(1);
            const Cx<T>* grow = J->green + (size_t)kyf * J->Hz * 6;
            const size_t gsx = (size_t)J->Hy * J->Hz * 6;
            for (int e = task; e < nw; e += NT) {
                const int kxf = (int)((unsigned)e / (unsigned)nkz), kzl = e - kxf * nkz;
                const int kz = t.f0 / 3 + kzl;
                const int kzf = (kz > Mz / 2) ? Mz - kz : kz;
                const T sz = (kz > Mz / 2) ? T(-1) : T(1);
                const Cx<T>* gp = grow + kxf * gsx + kzf * 6;
                const Cx<T> gxx = gld(gp), gyy = gld(gp + 3), gzz = gld(gp + 5);
                const Cx<T> gxy = cscale(gld(gp + 1), sy);
                const Cx<T> gxz = cscale(gld(gp + 2), sz);
                const Cx<T> gyz = cscale(gld(gp + 4), sy * sz);
                Cx<T>* b0 = A + (3 * kzl) * ps;
                green_apply(b0 + kxf, ps, gxx, gxy, gxz, gyy, gyz, gzz);
                if (kxf != 0 && 2 * kxf != LC) green_apply(b0 + (LC - kxf), ps, gxx, cneg(gxy), cneg(gxz), gyy, gyz, gzz);
            }
        }
        __syncthreads();
        // B': inverse DFT over k1, times conj(w_L^(n1 k2)), to the same slots
        if (actb) {
            // an opaque copy of k2: step A read the same twiddles tw[qa m],
            // and keeping them live from A to here would spill
            int k2o = k2;
            __asm__ volatile("" : "+v"(k2o));
            Cx<T>* cp = A + ca * ps + k2o;
            sfor<0, L1>([&](auto c1) { v[decltype(c1)::value] = cp[decltype(c1)::value * L2]; });
            sdft<T, L1, true>(v);
            cp[0] = v[0];
            sfor<1, L1>([&](auto c1) {
                constexpr int n1 = decltype(c1)::value;
                cp[n1 * L2] = cmulc(v[n1], tw[n1 * k2o]);
            });
        }
        __syncthreads();
        // A': inverse DFT over k2, store the x < Nx outputs
        if (qa < L1) {
            // opaque copies: the output addresses are recomputed rather than
            // kept live from step A
            int n1 = qa, cao = ca;
            __asm__ volatile("" : "+v"(n1), "+v"(cao));
            Cx<T> w[L2];
            const Cx<T>* row = A + cao * ps + n1 * L2;
            sfor<0, L2>([&](auto c2) { w[decltype(c2)::value] = row[decltype(c2)::value]; });
            sdft<T, L2, true>(w);
            sfor<0, H2>([&](auto c2) {
                constexpr int n2 = decltype(c2)::value;
                const int x = n1 + L1 * n2;
                if (x < Nx) gst(base + (x * sx + cao), w[n2]);
            });
        }
        __syncthreads();
    }
}

// P5: z-inverse, crop to the box, diagonal term and solver epilogues.
// Tile = pz consecutive (x, y) pencils.
template <typename T, int MODE, int LC>
__global__ void TDDA_PASS_BOUNDS k_p5(const JobDesc<T>* __restrict__ jobs, const int* __restrict__ tstart,
                                      int njobs, int diag)
{
    extern __shared__ __align__(16) char smem[];
    constexpr bool FT = LC > 0 && pass_full_twiddles<T>(0, LC);
    constexpr bool IP = LC > 0 && pz_inplace<T>(LC);  // one LDS buffer, in-place transforms
    __shared__ double red[16 * kPart];
    constexpr int EP = ept_len<T>(LC);
    const int ntot = tstart[njobs];
    Pre<T, EP> pr;
    ZTile cur;
    cur.live = false;
    for (int tl = blockIdx.x;; tl += gridDim.x) {
        prefetch_landed();
        if (cur.live) {
            const JobDesc<T>* J = jobs + cur.job;
            const int Mz = J->Mz, ps = J->ps_z;
            Cx<T>* twl = reinterpret_cast<Cx<T>*>(smem);
            Cx<T>* A = twl + pass_tw_slots(FT, Mz, J->plz.tb);
            load_pass_twiddles<T, FT>(twl, J->twz, J->twfz, J->plz);
            RowWalk w(Mz * 3, 1.0f / (float)(Mz * 3));
#pragma unroll
            for (int k = 0; k < EP; k++, w.next()) {
                if (w.row < cur.np) {
                    int kz, c;
                    split3(w.col, kz, c);
                    A[(w.row * 3 + c) * ps + lpad<T>(kz)] = pr.get(k);
                }
            }
        }
        ZTile nxt;
        nxt.live = false;
        if (tl < ntot) {
            nxt = z_tile(jobs, tstart, njobs, tl, 0, MODE != OUT_PLAIN);
            if (nxt.live) {
                const JobDesc<T>* J = jobs + nxt.job;
                const int Mz = J->Mz;
                const Cx<T>* w1 = J->w1 + (size_t)nxt.q0 * Mz * 3;
#pragma unroll
                for (int k = 0; k < EP; k++) {
                    const int e = threadIdx.x + k * kThreads;
                    if (e < nxt.np * Mz * 3) pr.set(k, gld(w1 + e));
                }
            }
        }
        if (cur.live) {
            const JobDesc<T>* J = jobs + cur.job;
            const int Nz = J->Nz, Mz = J->Mz, ps = J->ps_z;
            Cx<T>* twl = reinterpret_cast<Cx<T>*>(smem);
            Cx<T>* A = twl + pass_tw_slots(FT, Mz, J->plz.tb);
            Cx<T>* B = IP ? A : A + 3 * J->pz * ps;
            __syncthreads();
            const Cx<T>* res;
            if constexpr (IP) res = ct_fft_ip<T, LC, true, np_max<T>(LC), kPruneOut, FT>(A, cur.np * 3, twl);
            else res = pass_fft<T, LC, true, kPruneOut, FT>(A, B, cur.np * 3, ps, J->plz, twl);
            double acc[kPart] = {};
            // right Jacobi preconditioning: the operator is A D^-1, whose
            // diagonal term on v is v itself
            const bool prec = (MODE == OUT_AP || MODE == OUT_AT) && gstate(J->st)->prec != 0;
            const int nout = cur.np * Nz * 3;
            const int base = cur.q0 * Nz;
            RowWalk w(Nz * 3, 1.0f / (float)(Nz * 3));
            for (int e = threadIdx.x; e < nout; e += kThreads, w.next()) {
                const int pl = w.row;
                int z, c;
                split3(w.col, z, c);
                int i;
                if (J->cellmap) {
                    const int site = gldi(J->cellmap + base + pl * Nz + z);
                    if (site < 0) continue;
                    i = site * 3 + c;
                } else {
                    i = base * 3 + e;
                }
                const Cx<T> conv = res[(pl * 3 + c) * ps + lpad<T>(z)];
                if (MODE == OUT_PLAIN) {
                    Cx<T> y = conv;
                    if (diag) y = cadd(y, cmul(gld(J->al + i), gld(J->in + i)));
                    gst(J->out + i, y);
                } else if (MODE == OUT_RES) {
                    // r = E - A x ; r0 = r   (DDAModel.cpp:163-166)
                    const Cx<T> ax = cadd(conv, cmul(gld(J->al + i), gld(J->x + i)));
                    const Cx<T> ev = gld(J->e + i);
                    const Cx<T> r = csub(ev, ax);
                    gst(J->r + i, r);
                    gst(J->r0 + i, r);
                    const double2 rd = to_d(r), ed = to_d(ev);
                    acc[0] += rd.x * rd.x + rd.y * rd.y;
                    acc[1] += ed.x * ed.x + ed.y * ed.y;
                } else if (MODE == OUT_AP) {
                    const Cx<T> pv = gld(J->p + i);
                    const Cx<T> ap = cadd(conv, prec ? pv : cmul(gld(J->al + i), pv));
                    gst(J->ap + i, ap);
                    const double2 a = to_d(gld(J->r0 + i)), b = to_d(ap);
                    acc[0] += a.x * b.x + a.y * b.y;  // conj(r0) Ap
                    acc[1] += a.x * b.y - a.y * b.x;
                } else {  // OUT_AT
                    const Cx<T> tv = gld(J->t + i);
                    const Cx<T> at = cadd(conv, prec ? tv : cmul(gld(J->al + i), tv));
                    gst(J->at + i, at);
                    at_sums(acc, at, tv, gld(J->r0 + i));
                }
            }
            if (MODE != OUT_PLAIN) {
                block_sum<kPart>(acc, red);
                store_partials(J->part, cur.jt, acc);
            }
            __syncthreads();
        }
        if (tl >= ntot) break;
        cur = nxt;
    }
}

// ---------------------------------------------------------------------------
// Vector kernels of the solver (one tile = kThreads * kVecPer elements).
// ---------------------------------------------------------------------------
constexpr int kVecPer = 4;

// x0 = 0 start: r = r0 = E, x = 0 (A 0 = 0 exactly, so this equals
// DDAModel.cpp:163-166 with P = 0).
template <typename T>
__global__ void __launch_bounds__(kThreads) k_init_zero(const JobDesc<T>* __restrict__ jobs,
                                                        const int* __restrict__ tstart, int njobs)
{
    __shared__ double red[16 * kPart];
    const int job = find_job(tstart, njobs, blockIdx.x);
    const JobDesc<T>& J = jobs[job];
    if (job_inactive(J.st)) return;
    const int tile = blockIdx.x - tstart[job];
    const int n = J.N * 3;
    double acc[kPart] = {};
    const int i0 = tile * kThreads * kVecPer;
    for (int k = 0; k < kVecPer; k++) {
        const int i = i0 + k * kThreads + threadIdx.x;
        if (i < n) {
            const Cx<T> ev = gld(J.e + i);
            gst(J.r + i, ev);
            gst(J.r0 + i, ev);
            gst(J.x + i, Cx<T>{T(0), T(0)});
            const double2 d = to_d(ev);
            acc[0] += d.x * d.x + d.y * d.y;
        }
    }
    block_sum<kPart>(acc, red);
    store_partials(J.part, tile, acc);
}

// The pending update of a job that has stopped (converged, broken down or
// out of iterations): P += alpha p + z with z = zeta t, r = t - zeta At
// (DDAModel.cpp:226-230).  A running job's pending update is applied by its
// next P1; the finishing launch FIN_SETTLE then clears the flag.
template <typename T>
__global__ void __launch_bounds__(kThreads) k_settle(const JobDesc<T>* __restrict__ jobs,
                                                     const int* __restrict__ tstart, int njobs)
{
    const int job = find_job(tstart, njobs, blockIdx.x);
    const JobDesc<T>& J = jobs[job];
    const TDDA_GLOBAL SolverState* stc = gstate(J.st);
    if (stc->status == ST_RUNNING || !stc->pend) return;
    const int tile = blockIdx.x - tstart[job];
    const int n = J.N * 3;
    const Cx<T> al = from_d<T>(stc->alpha[0], stc->alpha[1]);
    const Cx<T> ze = from_d<T>(stc->zeta[0], stc->zeta[1]);
    const int i0 = tile * kThreads * kVecPer;
    for (int k = 0; k < kVecPer; k++) {
        const int i = i0 + k * kThreads + threadIdx.x;
        if (i < n) {
            const Cx<T> tv = gld(J.t + i);
            gst(J.r + i, csub(tv, cmul(ze, gld(J.at + i))));
            gst(J.x + i, cadd(gld(J.x + i), cadd(cmul(al, gld(J.p + i)), cmul(ze, tv))));
        }
    }
}

// Right Jacobi preconditioning: the iterate y of A D^-1 y = E becomes
// P = D^-1 y (DIV) once the job has stopped and settled; a solve from a
// given P starts from y = D P (!DIV).
template <typename T, bool DIV>
__global__ void __launch_bounds__(kThreads) k_prec_x(const JobDesc<T>* __restrict__ jobs,
                                                     const int* __restrict__ tstart, int njobs)
{
    const int job = find_job(tstart, njobs, blockIdx.x);
    const JobDesc<T>& J = jobs[job];
    const TDDA_GLOBAL SolverState* stc = gstate(J.st);
    if (!stc->prec || (DIV && stc->status == ST_RUNNING)) return;
    const int tile = blockIdx.x - tstart[job];
    const int n = J.N * 3;
    const int i0 = tile * kThreads * kVecPer;
    for (int k = 0; k < kVecPer; k++) {
        const int i = i0 + k * kThreads + threadIdx.x;
        if (i < n) {
            const Cx<T> xv = gld(J.x + i), a = gld(J.al + i);
            gst(J.x + i, DIV ? cdivv(xv, a) : cmul(xv, a));
        }
    }
}

// Totals of the partial sums of one reducing pass and the scalar step that
// follows it.  One workgroup per job; the per-tile partials are summed in a
// fixed order (thread-strided, then block_sum), so results do not depend on
// which other jobs share the launches.
enum { FIN_RES = 0, FIN_AP = 1, FIN_AT = 2, FIN_INIT = 3, FIN_SETTLE = 4 };

template <typename T, int KIND>
__global__ void __launch_bounds__(kThreads) k_finish(const JobDesc<T>* __restrict__ jobs, int njobs)
{
    __shared__ double red[16 * kPart];
    const JobDesc<T>& J = jobs[blockIdx.x];
    if (KIND == FIN_SETTLE) {
        if (threadIdx.x == 0 && J.st->status != ST_RUNNING) J.st->pend = 0;
        return;
    }
    if (job_inactive(J.st)) return;
    const int ntiles = KIND == FIN_INIT ? (int)((3 * (size_t)J.N + kThreads * kVecPer - 1) / (kThreads * kVecPer))
                                        : (J.Nx * J.Ny + J.pz - 1) / J.pz;
    double acc[kPart] = {};
    for (int i = threadIdx.x; i < ntiles; i += kThreads) {
#pragma unroll
        for (int k = 0; k < kPart; k++) acc[k] += *(const TDDA_GLOBAL double*)(J.part + i * kPart + k);
    }
    block_sum<kPart>(acc, red);
    if (threadIdx.x != 0) return;
    SolverState* st = J.st;
    const double* tot = acc;
    if (KIND == FIN_RES) {
        st->rho[0] = tot[0];  // r0.r = |r|^2
        st->rho[1] = 0.0;
        st->enorm = sqrt(tot[1]);
        st->rnorm = sqrt(tot[0]);
    } else if (KIND == FIN_AP) {
        // alpha = r0.r / r0.Ap   (DDAModel.cpp:168, 221)
        cdiv(st->rho[0], st->rho[1], tot[0], tot[1], st->alpha[0], st->alpha[1]);
    } else if (KIND == FIN_INIT) {
        st->rho[0] = tot[0];
        st->rho[1] = 0.0;
        st->enorm = sqrt(tot[0]);
        st->rnorm = sqrt(tot[0]);
    } else {  // FIN_AT
        // zeta = At.t / At.At   (DDAModel.cpp:172, 225)
        double zr, zi;
        cdiv(tot[0], tot[1], tot[2], 0.0, zr, zi);
        st->zeta[0] = zr;
        st->zeta[1] = zi;
        // r = t - zeta At (DDAModel.cpp:230), pending: rl = r, then
        // r0.r = r0.t - zeta r0.At and
        // |r|^2 = |t|^2 - 2 Re(conj(zeta) At.t) + |zeta|^2 At.At
        st->pend = 1;
        st->rho_prev[0] = st->rho[0];
        st->rho_prev[1] = st->rho[1];
        st->rho[0] = tot[4] - (zr * tot[6] - zi * tot[7]);
        st->rho[1] = tot[5] - (zr * tot[7] + zi * tot[6]);
        const double rr = tot[3] - 2.0 * (zr * tot[0] + zi * tot[1]) + (zr * zr + zi * zi) * tot[2];
        st->rnorm = sqrt(fmax(rr, 0.0));
        st->err = st->rnorm / st->enorm;
        if (st->first) {
            st->first = 0;
        } else {
            st->iter += 1;
            if (st->err <= st->tol) {
                st->status = ST_CONVERGED;
                return;
            }
        }
        if (st->iter >= st->maxit) {
            st->status = ST_MAXIT;
            return;
        }
        if (st->rho_prev[0] == 0.0 && st->rho_prev[1] == 0.0) {
            st->status = ST_BREAKDOWN;
            st->iter += 1;
        }
    }
}

}  // namespace tdda
