// tdda_setup.hpp -- host-library-only kernels: precision conversion for the
// mixed solver, sparse entry access, and the operator set-up (interaction
// tensor on the padded grid, its 3D FFT, octant extraction).  Included by
// tdda_capi.hip only.
#pragma once
#include "tdda_fcd.hpp"
#include "tdda_kernels.hpp"

namespace tdda {

// ---------------------------------------------------------------------------
// Mixed precision glue: inner right-hand side and solution accumulation.
// ---------------------------------------------------------------------------
// e32 = c64(r64)
static __global__ void __launch_bounds__(kThreads) k_demote(const Cx<double>* __restrict__ src, Cx<float>* __restrict__ dst,
                                                     size_t n)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        const Cx<double> v = src[i];
        dst[i] = Cx<float>{float(v.x), float(v.y)};
    }
}

// x64 += c128(x32), or x64 += c128(x32) / al64 when the inner solve ran
// with right Jacobi preconditioning (al64 not null: x32 is y of A D^-1 y = r)
static __global__ void __launch_bounds__(kThreads) k_accum(Cx<double>* __restrict__ x, const Cx<float>* __restrict__ d,
                                                    size_t n, const Cx<double>* __restrict__ al = nullptr)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        const Cx<float> v = d[i];
        Cx<double> dv{double(v.x), double(v.y)};
        if (al) dv = cdivv(dv, al[i]);
        Cx<double> a = x[i];
        a.x += dv.x;
        a.y += dv.y;
        x[i] = a;
    }
}

// The same glue over a whole job set in one launch (the vector tiles of
// the set's tables, which share one tiling: it depends on N only).
//   MIX_DEMOTE  e32 = c64(r of the residual table)
//   MIX_ACCUM   x64 += c128(x32), divided by al64 with right Jacobi
//               preconditioning
// Both skip the jobs whose complex64 state is idle (not part of this
// refinement cycle).
enum { MIX_DEMOTE = 1, MIX_ACCUM = 2 };

template <int OP>
static __global__ void __launch_bounds__(kThreads) k_mix(const JobDesc<double>* __restrict__ j64,
                                                         const JobDesc<double>* __restrict__ jo,
                                                         const JobDesc<float>* __restrict__ j32,
                                                         const int* __restrict__ tv, int njobs, int prec)
{
    const int job = find_job(tv, njobs, blockIdx.x);
    if (gstate(j32[job].st)->idle) return;
    const int tile = blockIdx.x - tv[job];
    const int n = j64[job].N * 3;
    const int i0 = tile * kThreads * kVecPer;
    for (int k = 0; k < kVecPer; k++) {
        const int i = i0 + k * kThreads + threadIdx.x;
        if (i >= n) break;
        if (OP == MIX_DEMOTE) {
            const Cx<double> v = gld(jo[job].r + i);
            gst(const_cast<Cx<float>*>(j32[job].e) + i, Cx<float>{float(v.x), float(v.y)});
        } else {
            const Cx<float> v = gld(j32[job].x + i);
            Cx<double> dv{double(v.x), double(v.y)};
            if (prec) dv = cdivv(dv, gld(j64[job].al + i));
            Cx<double> a = gld(j64[job].x + i);
            a.x += dv.x;
            a.y += dv.y;
            gst(j64[job].x + i, a);
        }
    }
}

// Sparse entry access: out[k] = v[idx[k]] and v[idx[k]] = vals[k].
static __global__ void __launch_bounds__(kThreads) k_gather(const Cx<double>* __restrict__ v, const int* __restrict__ idx,
                                                     int n, Cx<double>* __restrict__ out)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) out[k] = v[idx[k]];
}

static __global__ void __launch_bounds__(kThreads) k_scatter(Cx<double>* __restrict__ v, const int* __restrict__ idx, int n,
                                                      const Cx<double>* __restrict__ vals)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) v[idx[k]] = vals[k];
}

// ---------------------------------------------------------------------------
// Green spectrum set-up (AProductCore.cpp:54-229, built on the device).
// ---------------------------------------------------------------------------
// Interaction tensor AProductCore.cpp:588-623 (LDR_inter).
__device__ __forceinline__ void ldr_inter(double K, double x, double y, double z, Cx<double>* out)
{
    const double xs = x * x, ys = y * y, zs = z * z;
    const double rn = sqrt(xs + ys + zs);
    if (rn == 0.0) {
        for (int l = 0; l < 6; l++) out[l] = Cx<double>{0.0, 0.0};
        return;
    }
    const double r2 = rn * rn, r3 = rn * rn * rn;
    double sn, cs;
    sincos(K * rn, &sn, &cs);
    const Cx<double> c1{cs / r3, sn / r3};
    const Cx<double> c2{1.0 / r2, -K * rn / r2};
    const Cx<double> c3{K * K - 3.0 * c2.x, -3.0 * c2.y};
    const double kk = K * K;
    Cx<double> a[6];
    a[0] = Cx<double>{c2.x * (r2 - 3.0 * xs) - kk * (ys + zs), c2.y * (r2 - 3.0 * xs)};
    a[1] = cscale(c3, x * y);
    a[2] = cscale(c3, z * x);
    a[3] = Cx<double>{c2.x * (r2 - 3.0 * ys) - kk * (zs + xs), c2.y * (r2 - 3.0 * ys)};
    a[4] = cscale(c3, y * z);
    a[5] = Cx<double>{c2.x * (r2 - 3.0 * zs) - kk * (xs + ys), c2.y * (r2 - 3.0 * zs)};
    for (int l = 0; l < 6; l++) out[l] = cmul(c1, a[l]);
}

// Fills the 6-component tensor on the whole padded grid [Mx][My][Mz][6]:
// index i < N -> offset i d; i >= M - (N - 1) -> (i - M) d; otherwise zero
// (AProductCore.cpp:54-194 with 2N-1 replaced by M).  With pm >= 0 the
// tensor of a structure periodic in x and y (AProductCore.cpp:251-536):
// the sum over the images m = -pm..pm, n = -pn..pn of the tensor at
// (x + m Lm, y + n Ln, z), each damped by exp(-(gamma K r)^4), gamma = 0.01,
// r the image's distance (A_dic_generator, AProductCore.cpp:573-586); the
// (0, 0) image of offset 0 is zero like the tensor itself.  fcd != 0 selects
// the FCD tensor instead of LDR.
// The tensor of one offset: LDR (AProductCore.cpp:588-623) or, fcd != 0,
// FCD (AProductCore.cpp:625-695); A_dic_generator, AProductCore.cpp:560-571.
__device__ __forceinline__ void inter_tensor(int fcd, double K, double d, double x, double y, double z,
                                             Cx<double>* out)
{
    if (fcd) {
        double t[12];
        fcd_inter(K, d, x, y, z, t);
        for (int l = 0; l < 6; l++) out[l] = Cx<double>{t[2 * l], t[2 * l + 1]};
    } else {
        ldr_inter(K, x, y, z, out);
    }
}

static __global__ void __launch_bounds__(kThreads) k_green_fill(Cx<double>* __restrict__ g, int Nx, int Ny, int Nz, int Mx,
                                                         int My, int Mz, double d, double K, int pm, int pn,
                                                         double Lm, double Ln, int fcd)
{
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t M = (size_t)Mx * My * Mz;
    if (idx >= M) return;
    const int k = (int)(idx % Mz);
    const int j = (int)((idx / Mz) % My);
    const int i = (int)(idx / ((size_t)Mz * My));
    const bool okx = (i <= Nx - 1) || (i >= Mx - (Nx - 1));
    const bool oky = (j <= Ny - 1) || (j >= My - (Ny - 1));
    const bool okz = (k <= Nz - 1) || (k >= Mz - (Nz - 1));
    Cx<double> a[6];
    if (okx && oky && okz) {
        const int oi = (i <= Nx - 1) ? i : i - Mx;
        const int oj = (j <= Ny - 1) ? j : j - My;
        const int ok = (k <= Nz - 1) ? k : k - Mz;
        if (pm < 0) {
            inter_tensor(fcd, K, d, d * oi, d * oj, d * ok, a);
        } else {
            const double gamma = 0.01;
            for (int l = 0; l < 6; l++) a[l] = Cx<double>{0.0, 0.0};
            for (int m = -pm; m <= pm; m++) {
                for (int n = -pn; n <= pn; n++) {
                    const double x = d * oi + m * Lm, y = d * oj + n * Ln, z = d * ok;
                    Cx<double> t[6];
                    inter_tensor(fcd, K, d, x, y, z, t);
                    const double damp = exp(-pow(gamma * K * sqrt(x * x + y * y + z * z), 4));
                    for (int l = 0; l < 6; l++) a[l] = cadd(a[l], cscale(t[l], damp));
                }
            }
        }
    } else {
        for (int l = 0; l < 6; l++) a[l] = Cx<double>{0.0, 0.0};
    }
    for (int l = 0; l < 6; l++) g[idx * 6 + l] = a[l];
}

// Generic in-place transform along one axis of an array [D0][D1][D2][NC]
// viewed as pencils: element (o, a, f) at o*so + a*sa + f, f in [0, F).
// A tile is (o, tf consecutive f).  Used for the Green spectrum (NC = 6).
template <typename T, bool INV>
__global__ void __launch_bounds__(kThreads) k_fft_strided(Cx<T>* __restrict__ data, int L, size_t so, size_t sa,
                                                          int F, int tf, int ntf, int ps, FftPlan pl,
                                                          const Cx<T>* __restrict__ tw)
{
    extern __shared__ __align__(16) char smem[];
    const int o = blockIdx.x / ntf;
    const int f0 = (blockIdx.x - o * ntf) * tf;
    const int nf = min(tf, F - f0);
    Cx<T>* A = reinterpret_cast<Cx<T>*>(smem);
    Cx<T>* B = A + tf * ps;
    Cx<T>* base = data + (size_t)o * so + f0;
    for (int e = threadIdx.x; e < L * nf; e += blockDim.x) {
        const int a = e / nf;
        const int f = e - a * nf;
        A[f * ps + a] = base[(size_t)a * sa + f];
    }
    __syncthreads();
    const Cx<T>* res = lds_fft<T, INV>(A, B, nf, ps, pl, tw);
    for (int e = threadIdx.x; e < L * nf; e += blockDim.x) {
        const int a = e / nf;
        const int f = e - a * nf;
        base[(size_t)a * sa + f] = res[f * ps + a];
    }
}

// Contiguous-axis variant: pencil q occupies data[q*L*NC ... ) as [a][NC].
template <typename T, bool INV>
__global__ void __launch_bounds__(kThreads) k_fft_contig(Cx<T>* __restrict__ data, int Q, int L, int NC, int P, int ps,
                                                         FftPlan pl, const Cx<T>* __restrict__ tw)
{
    extern __shared__ __align__(16) char smem[];
    const int q0 = blockIdx.x * P;
    const int np = min(P, Q - q0);
    Cx<T>* A = reinterpret_cast<Cx<T>*>(smem);
    Cx<T>* B = A + P * NC * ps;
    Cx<T>* base = data + (size_t)q0 * L * NC;
    const int n = np * L * NC;
    for (int e = threadIdx.x; e < n; e += blockDim.x) {
        const int pl_ = e / (L * NC);
        const int rem = e - pl_ * (L * NC);
        const int a = rem / NC;
        const int c = rem - a * NC;
        A[(pl_ * NC + c) * ps + a] = base[e];
    }
    __syncthreads();
    const Cx<T>* res = lds_fft<T, INV>(A, B, np * NC, ps, pl, tw);
    for (int e = threadIdx.x; e < n; e += blockDim.x) {
        const int pl_ = e / (L * NC);
        const int rem = e - pl_ * (L * NC);
        const int a = rem / NC;
        const int c = rem - a * NC;
        base[e] = res[(pl_ * NC + c) * ps + a];
    }
}

// Octant extraction with the 1/M scale folded in: full [Mx][My][Mz][6] FP64
// spectrum -> [Hx][Hy][Hz][6] of type T.
template <typename T>
__global__ void __launch_bounds__(kThreads) k_green_octant(const Cx<double>* __restrict__ full, Cx<T>* __restrict__ oct,
                                                           int Mx, int My, int Mz, int Hx, int Hy, int Hz, double scale)
{
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t H = (size_t)Hx * Hy * Hz * 6;
    if (idx >= H) return;
    const int l = (int)(idx % 6);
    const size_t pnt = idx / 6;
    const int k = (int)(pnt % Hz);
    const int j = (int)((pnt / Hz) % Hy);
    const int i = (int)(pnt / ((size_t)Hz * Hy));
    const Cx<double> v = full[(((size_t)i * My + j) * Mz + k) * 6 + l];
    oct[idx] = Cx<T>{T(v.x * scale), T(v.y * scale)};
}

// ---------------------------------------------------------------------------
// Request set-up on the device (request loop).
// ---------------------------------------------------------------------------
// dst[slot[i]] = src[i]: solver states of the jobs that start a phase.
static __global__ void __launch_bounds__(kThreads) k_put_states(SolverState* __restrict__ dst,
                                                         const int* __restrict__ slot,
                                                         const SolverState* __restrict__ src, int n)
{
    const int i = blockIdx.x * kThreads + threadIdx.x;
    if (i < n) dst[slot[i]] = src[i];
}

__device__ __forceinline__ Cx<double> cdiv(Cx<double> a, Cx<double> b)
{
    const double q = b.x * b.x + b.y * b.y;
    return Cx<double>{(a.x * b.x + a.y * b.y) / q, (a.y * b.x - a.x * b.y) / q};
}

// eps from a design value t by the material interpolation (DDAModel.cpp:54-61):
// mat[floor t] + (t - floor t) (mat[next] - mat[floor t]), next = floor t + 1
// for floor t = 0, else floor t.
__device__ __forceinline__ Cx<double> interp_eps(double tv, const Cx<double>* __restrict__ mat, int nmat)
{
    const int lf = (int)floor(tv);
    const int ln = lf >= 1 ? lf : lf + 1;
    const Cx<double> mf = mat[lf], mn = mat[min(ln, nmat - 1)];
    return cadd(mf, cscale(csub(mn, mf), tv - (double)lf));
}

// 1 / alpha_LDR (Get_Alpha, tools.cpp:617-630; DDAModel.cpp:62) for one
// incidence q = {nK[3], nE[3], ...}.
__device__ __forceinline__ Cx<double> ldr_inv_alpha(Cx<double> eps, const double* q, double K, double d)
{
    const double b1 = -1.891531, b2 = 0.1648469, b3 = -1.7700004;
    const double d3 = d * d * d, kd = K * d;
    const Cx<double> a1 =
        cscale(cdiv(Cx<double>{eps.x - 1.0, eps.y}, Cx<double>{eps.x + 2.0, eps.y}), 3.0 * d3 / (4.0 * M_PI));
    const double S = (q[3] * q[0]) * (q[3] * q[0]) + (q[4] * q[1]) * (q[4] * q[1]) + (q[5] * q[2]) * (q[5] * q[2]);
    const Cx<double> poly = Cx<double>{b1 + eps.x * b2 + eps.x * b3 * S, eps.y * b2 + eps.y * b3 * S};
    Cx<double> inner = cscale(poly, kd * kd);
    inner.y -= (2.0 / 3.0) * kd * kd * kd;
    const Cx<double> res = cadd(Cx<double>{1.0, 0.0}, cmul(cscale(a1, 1.0 / d3), inner));
    return cdiv(res, a1);  // 1 / (a1 / res)
}

// The DDAModel constructor per right-hand side (DDAModel.cpp:12-69):
// eps from the design value t (interp_eps), al = 1 / alpha_LDR and, when
// `plane`, the incident plane wave E = E0 nE exp(i K d nK . R)
// (DDAModel.cpp:47-51); otherwise E = 0.  inc holds 7 doubles per
// right-hand side: nK[3], nE[3], E0.  R == nullptr stands for the full box in
// build_a_bulk order.  Design values must lie in the material table (checked
// on the host).
static __global__ void __launch_bounds__(kThreads) k_request_fields(
    int N, const int* __restrict__ R, int Ny, int Nz, const double* __restrict__ t, int nmat,
    const Cx<double>* __restrict__ mat, double K, double d, const double* __restrict__ inc, int nrhs, int plane,
    Cx<double>* __restrict__ al, Cx<double>* __restrict__ E)
{
    const size_t n3 = 3 * (size_t)N;
    const size_t total = n3 * (size_t)nrhs;
    for (size_t g = (size_t)blockIdx.x * kThreads + threadIdx.x; g < total; g += (size_t)gridDim.x * kThreads) {
        const size_t j = g / n3;
        const size_t e = g - j * n3;
        const int i = (int)(e / 3), c = (int)(e - 3 * (size_t)i);
        const double* q = inc + 7 * j;
        al[g] = ldr_inv_alpha(interp_eps(t[e], mat, nmat), q, K, d);
        Cx<double> ev{0.0, 0.0};
        if (plane) {
            int x, y, z;
            if (R) {
                x = R[3 * i];
                y = R[3 * i + 1];
                z = R[3 * i + 2];
            } else {
                x = i / (Ny * Nz);
                const int r = i - x * Ny * Nz;
                y = r / Nz;
                z = r - y * Nz;
            }
            const double ph = K * d * (q[0] * x + q[1] * y + q[2] * z);
            double sn, cs;
            sincos(ph, &sn, &cs);
            ev = Cx<double>{q[6] * q[3 + c] * cs, q[6] * q[3 + c] * sn};
        }
        E[g] = ev;
    }
}

// Design update (CoreStructure::UpdateStr, CoreStructure.cpp:31-104) over any
// parameterization: free parameter i is Para index fi[i].  With a filter,
// Para_origin += step, clamped to [0, 1] (CoreStructure.cpp:48-56); without,
// Para itself.
static __global__ void __launch_bounds__(kThreads) k_design_step(double* __restrict__ v,
                                                              const int* __restrict__ fi,
                                                              const double* __restrict__ step, int n)
{
    const int i = blockIdx.x * kThreads + threadIdx.x;
    if (i < n) {
        double x = v[fi[i]] + step[i];
        x = x >= 1.0 ? 1.0 : x;
        x = x <= 0.0 ? 0.0 : x;
        v[fi[i]] = x;
    }
}

// Cone filter of free parameter i: its weights (w, Para index) in
// SpacePara's FreeWeight order (SpacePara.cpp:1232-1380), sum w origin /
// sum w in that order (CoreStructure.cpp:60-80), then SmoothDensity
// (FilterOption.cpp:49-60).  Contraction is off so that the sums round like
// the reference's.
static __global__ void __launch_bounds__(kThreads) k_design_filter(const double* __restrict__ origin,
                                                                const int* __restrict__ fi,
                                                                const int* __restrict__ ws,
                                                                const int* __restrict__ wp,
                                                                const double* __restrict__ wv, int n, double beta,
                                                                double eta, double* __restrict__ filtered,
                                                                double* __restrict__ para)
{
#pragma clang fp contract(off)
    const int i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= n) return;
    double num = 0.0, den = 0.0;
    for (int k = ws[i]; k < ws[i + 1]; k++) {
        const double w = wv[k];
        num += w * origin[wp[k]];
        den += w;
    }
    const double f = num / den;
    filtered[fi[i]] = f;
    double p;
    if (f <= eta) p = eta * (exp(-beta * (1 - f / eta)) - (1 - f / eta) * exp(-beta));
    else p = (1 - eta) * (1 - exp(-beta * (f - eta) / (1 - eta)) + (f - eta) / (1 - eta) * exp(-beta)) + eta;
    para[fi[i]] = p;
}

// al of every right-hand side from the parameter of each site
// (DDAModel::UpdateAlpha, DDAModel.cpp:470-484; diel_old of site i is
// Para(geometryPara(i)), CoreStructure.cpp:98-104).  inc: 6 doubles per
// right-hand side (nK[3], nE[3]).
static __global__ void __launch_bounds__(kThreads) k_design_alpha(const double* __restrict__ para,
                                                               const int* __restrict__ site_param, int N,
                                                               const Cx<double>* __restrict__ mat, int nmat, double K,
                                                               double d, const double* __restrict__ inc, int nrhs,
                                                               Cx<double>* __restrict__ al)
{
    const size_t n3 = 3 * (size_t)N;
    const size_t total = n3 * (size_t)nrhs;
    for (size_t g = (size_t)blockIdx.x * kThreads + threadIdx.x; g < total; g += (size_t)gridDim.x * kThreads) {
        const size_t j = g / n3;
        const size_t e = g - j * n3;
        const int i = (int)(e / 3);
        al[g] = ldr_inv_alpha(interp_eps(para[site_param[i]], mat, nmat), inc + 6 * j, K, d);
    }
}

// ---------------------------------------------------------------------------
// Objectives of the reference (ObjDDAModel.cpp) and their adjoint sources.
// The reference estimates the adjoint source 2 dF/dP of an objective F by a
// one-sided finite difference per entry of P, real and imaginary
// (EvoDDAModel::devp, EvoDDAModel.cpp:353-391):
//   devp_i = (F(P + e_i eps) - F) / eps + (F(P + i e_i eps) - F) / (i eps)
//          ~ dF/dRe P_i - i dF/dIm P_i.
// Here it is the same quantity in closed form.  Site coordinates come from R
// or, for R == nullptr, the full box in build_a_bulk order.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void site_xyz(const int* __restrict__ R, int i, int Ny, int Nz, int& x, int& y, int& z)
{
    if (R) {
        x = R[3 * i];
        y = R[3 * i + 1];
        z = R[3 * i + 2];
    } else {
        x = i / (Ny * Nz);
        const int r = i - x * Ny * Nz;
        y = r / Nz;
        z = r - y * Nz;
    }
}

// ObjPointEDDAModel (ObjDDAModel.cpp:20-97): the field at r0 is
// E = E_ext - sum_i A(r0 - r_i d) P_i; per workgroup the partial sum of
// -A(r0 - r_i d) P_i (3 complex values) in part[6 * blockIdx.x ...].
static __global__ void __launch_bounds__(kThreads) k_obj_point_e(const Cx<double>* __restrict__ P, const int* __restrict__ R,
                                                              int N, int Ny, int Nz, double d, double K, int fcd,
                                                              double x0, double y0, double z0,
                                                              double* __restrict__ part)
{
    __shared__ double red[16 * 6];
    double acc[6] = {0, 0, 0, 0, 0, 0};
    for (int i = blockIdx.x * kThreads + threadIdx.x; i < N; i += gridDim.x * kThreads) {
        int x, y, z;
        site_xyz(R, i, Ny, Nz, x, y, z);
        Cx<double> a[6];
        inter_tensor(fcd, K, d, x0 - d * x, y0 - d * y, z0 - d * z, a);
        const Cx<double> p0 = P[3 * i], p1 = P[3 * i + 1], p2 = P[3 * i + 2];
        const Cx<double> e0 = cadd(cadd(cmul(a[0], p0), cmul(a[1], p1)), cmul(a[2], p2));
        const Cx<double> e1 = cadd(cadd(cmul(a[1], p0), cmul(a[3], p1)), cmul(a[4], p2));
        const Cx<double> e2 = cadd(cadd(cmul(a[2], p0), cmul(a[4], p1)), cmul(a[5], p2));
        acc[0] -= e0.x;
        acc[1] -= e0.y;
        acc[2] -= e1.x;
        acc[3] -= e1.y;
        acc[4] -= e2.x;
        acc[5] -= e2.y;
    }
#pragma unroll
    for (int k = 0; k < 6; k++) {
        double v = acc[k];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
        acc[k] = v;
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0)
        for (int k = 0; k < 6; k++) red[wid * 6 + k] = acc[k];
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 0; k < 6; k++) {
            double v = 0.0;
            for (int w = 0; w < kThreads / 64; w++) v += red[w * 6 + k];
            part[blockIdx.x * 6 + k] = v;
        }
    }
}

// Adjoint source of PointE, F = |E|: 2 dF/dP_ic = -(1/F) sum_a A_ac conj(E_a).
static __global__ void __launch_bounds__(kThreads) k_obj_point_e_source(const int* __restrict__ R, int N, int Ny,
                                                                     int Nz, double d, double K, int fcd, double x0,
                                                                     double y0, double z0, Cx<double> E0, Cx<double> E1,
                                                                     Cx<double> E2, double invF,
                                                                     Cx<double>* __restrict__ dst)
{
    for (int i = blockIdx.x * kThreads + threadIdx.x; i < N; i += gridDim.x * kThreads) {
        int x, y, z;
        site_xyz(R, i, Ny, Nz, x, y, z);
        Cx<double> a[6];
        inter_tensor(fcd, K, d, x0 - d * x, y0 - d * y, z0 - d * z, a);
        const int col[3][3] = {{0, 1, 2}, {1, 3, 4}, {2, 4, 5}};
        for (int c = 0; c < 3; c++) {
            // sum_a A_ac conj(E_a)
            Cx<double> s = cmulc(a[col[0][c]], E0);
            s = cadd(s, cmulc(a[col[1][c]], E1));
            s = cadd(s, cmulc(a[col[2][c]], E2));
            dst[3 * i + c] = cscale(s, -invF);
        }
    }
}

// DDAModel::solve_E (DDAModel.cpp:530-565): the field at lattice points
// pts (units of d), one workgroup per point:
// E = E_ext - sum_i A(d pt - d R_i) P_i, with the incident field as the
// reference writes it -- component c's cosine takes K (nK_x q_c + nK_y y +
// nK_z z) with q = (x, y, z), its sine K nK . r (DDAModel.cpp:542-544).
static __global__ void __launch_bounds__(kThreads) k_field_points(const Cx<double>* __restrict__ P,
                                                               const int* __restrict__ R, int N, int Ny, int Nz,
                                                               double d, double K, int fcd,
                                                               const int* __restrict__ pts, double E0, double nKx,
                                                               double nKy, double nKz, double nEx, double nEy,
                                                               double nEz, Cx<double>* __restrict__ out)
{
    __shared__ double red[(kThreads / 64) * 6];
    const int q = blockIdx.x;
    const double x0 = d * pts[3 * q], y0 = d * pts[3 * q + 1], z0 = d * pts[3 * q + 2];
    double acc[6] = {0, 0, 0, 0, 0, 0};
    for (int i = threadIdx.x; i < N; i += kThreads) {
        int x, y, z;
        site_xyz(R, i, Ny, Nz, x, y, z);
        Cx<double> a[6];
        inter_tensor(fcd, K, d, x0 - d * x, y0 - d * y, z0 - d * z, a);
        const Cx<double> p0 = P[3 * i], p1 = P[3 * i + 1], p2 = P[3 * i + 2];
        const Cx<double> e0 = cadd(cadd(cmul(a[0], p0), cmul(a[1], p1)), cmul(a[2], p2));
        const Cx<double> e1 = cadd(cadd(cmul(a[1], p0), cmul(a[3], p1)), cmul(a[4], p2));
        const Cx<double> e2 = cadd(cadd(cmul(a[2], p0), cmul(a[4], p1)), cmul(a[5], p2));
        acc[0] += e0.x;
        acc[1] += e0.y;
        acc[2] += e1.x;
        acc[3] += e1.y;
        acc[4] += e2.x;
        acc[5] += e2.y;
    }
#pragma unroll
    for (int k = 0; k < 6; k++) {
        double v = acc[k];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
        acc[k] = v;
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0)
        for (int k = 0; k < 6; k++) red[wid * 6 + k] = acc[k];
    __syncthreads();
    if (threadIdx.x == 0) {
        double sum[6];
        for (int k = 0; k < 6; k++) {
            double v = 0.0;
            for (int w = 0; w < kThreads / 64; w++) v += red[w * 6 + k];
            sum[k] = v;
        }
        const double ph = K * (nKx * x0 + nKy * y0 + nKz * z0);
        const double sn = sin(ph);
        const double cs[3] = {cos(ph), cos(K * (nKx * y0 + nKy * y0 + nKz * z0)),
                              cos(K * (nKx * z0 + nKy * y0 + nKz * z0))};
        const double ne[3] = {nEx, nEy, nEz};
        for (int c = 0; c < 3; c++)
            out[3 * q + c] = Cx<double>{E0 * ne[c] * cs[c] - sum[2 * c], E0 * ne[c] * sn - sum[2 * c + 1]};
    }
}

// IntegratedE's region test as the reference writes it,
// (lo <= v <= hi) with C++ semantics: (lo <= v) is 0 or 1, compared with hi
// (ObjDDAModel.cpp:138).
__device__ __forceinline__ bool chained_le(int lo, int v, int hi) { return (lo <= v ? 1 : 0) <= hi; }

__device__ __forceinline__ double smooth_density_dev(double x, double eta, double beta)
{
    // ObjDDAModel.cpp:5-16 (same function as FilterOption::SmoothDensity)
    if (x <= eta) return eta * (exp(-beta * (1 - x / eta)) - (1 - x / eta) * exp(-beta));
    return (1 - eta) * (1 - exp(-beta * (x - eta) / (1 - eta)) + (x - eta) / (1 - eta) * exp(-beta)) + eta;
}

// ObjIntegratedEDDAModel (ObjDDAModel.cpp:100-177): E_int = sum over sites
// in the region of SmoothDensity(diel_old_3i) sum_c |al_c P_c|^pow; per
// workgroup partial sums in part[blockIdx.x].  With `src`, also the
// adjoint source of F = log E_int up to the factor 1 / E_int:
// pow f |al|^pow |P|^(pow - 2) conj(P) (zero where P = 0).
static __global__ void __launch_bounds__(kThreads) k_obj_integrated_e(
    const Cx<double>* __restrict__ P, const Cx<double>* __restrict__ al, const double* __restrict__ diel_old,
    const int* __restrict__ R, int N, int Ny, int Nz, int pw, int xmin, int xmax, int ymin, int ymax, int zmin,
    int zmax, double eta, double beta, double* __restrict__ part, Cx<double>* __restrict__ src)
{
    __shared__ double red[16];
    double acc = 0.0;
    for (int i = blockIdx.x * kThreads + threadIdx.x; i < N; i += gridDim.x * kThreads) {
        int x, y, z;
        site_xyz(R, i, Ny, Nz, x, y, z);
        const bool in = chained_le(xmin, x, xmax) && chained_le(ymin, y, ymax) && chained_le(zmin, z, zmax);
        const double f = in ? smooth_density_dev(diel_old[3 * i], eta, beta) : 0.0;
        for (int c = 0; c < 3; c++) {
            const Cx<double> a = al[3 * i + c], p = P[3 * i + c];
            const double aa = sqrt(a.x * a.x + a.y * a.y), pp = sqrt(p.x * p.x + p.y * p.y);
            if (in) acc += f * pow(aa * pp, (double)pw);
            if (src) {
                Cx<double> g{0.0, 0.0};
                if (in && pp > 0.0) g = cscale(Cx<double>{p.x, -p.y}, pw * f * pow(aa, (double)pw) * pow(pp, (double)(pw - 2)));
                src[3 * i + c] = g;
            }
        }
    }
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) red[wid] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double v = 0.0;
        for (int w = 0; w < kThreads / 64; w++) v += red[w];
        part[blockIdx.x] = v;
    }
}

// dst *= s (the 1 / E_int of the IntegratedE source)
static __global__ void __launch_bounds__(kThreads) k_scale(Cx<double>* __restrict__ v, size_t n, double s)
{
    const size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x;
    if (i < n) v[i] = cscale(v[i], s);
}

// Sparse right-hand sides (adjoint sources): E[j][idx[j][k]] = val[j][k].
static __global__ void __launch_bounds__(kThreads) k_sparse_rhs(Cx<double>* __restrict__ E, size_t n3, int nrhs,
                                                         int nsrc, const int* __restrict__ idx,
                                                         const Cx<double>* __restrict__ val)
{
    const int g = blockIdx.x * kThreads + threadIdx.x;
    if (g < nrhs * nsrc) {
        const int j = g / nsrc;
        E[(size_t)j * n3 + idx[g]] = val[g];
    }
}

}  // namespace tdda
