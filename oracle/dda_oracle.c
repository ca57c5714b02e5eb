/*
 * dda_oracle.c -- CPU FP64 restatement of the reference's DDA operator and
 * BiCGSTAB solve (Topo-DDA-forWin64).
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: it is loaded by the
 * tests, by __graft_entry__.smoke() and by the cpu_baseline leg of bench.py.
 * The product (libtopodda_hip.so and the host classes) never links, loads or
 * calls it.
 *
 * Every function cites the reference file:line it restates.  Paths are
 * relative to the reference's Topo-DDA-forWin64/ directory.
 *
 * Complex vectors cross the C boundary as interleaved (re, im) doubles.
 * Site vectors are site-major with the three Cartesian components
 * interleaved: entry 3*i+c is component c of dipole i (AProductCore.cpp:705-714).
 *
 * Parity pins (each enforced in tests/test_oracle.py):
 *   orc_ldr_inter      closed-form dyadic Green function -G (independent form),
 *                      trace = -2 k^2 e^{ikr}/r, A r^ = 2(ikr-1)e^{ikr}/r^3 r^,
 *                      long-double evaluation.
 *   orc_get_alpha      the reference's own Get_Alpha (tools.cpp compiled into
 *                      oracle/_ref) and golden vectors made from it.
 *   orc_einc           closed form |E| = E0 |n_E|, phase = K d n_K.R.
 *   orc_diel_al        special cases t=0, t=1, midpoint.
 *   orc_aproduct_*     O(N^2) direct sum in long double (tiny lattices), dense
 *                      assembly, numpy fftn, grid-size independence.
 *   orc_bicgstab       dense numpy.linalg.solve of the assembled 3N x 3N system.
 */
#include <complex.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef double complex cplx;

#define ORC_PI 3.14159265358979323846264338327950288L

/* ------------------------------------------------------------------------- */
/* Interaction tensor.  AProductCore.cpp:588-623 (LDR_inter), reached through */
/* A_dic_generator (AProductCore.cpp:560-571).  Output order xx xy xz yy yz zz */
/* (AProductCore.cpp:58-59), 6 complex values as 12 doubles.                   */
/* ------------------------------------------------------------------------- */
static void ldr_inter(double K, double x, double y, double z, cplx out[6])
{
    double xs = x * x, ys = y * y, zs = z * z;
    double rn = sqrt(xs + ys + zs);
    if (rn == 0.0) { /* zero self term, AProductCore.cpp:593-598 */
        for (int l = 0; l < 6; l++) out[l] = 0.0;
        return;
    }
    double r2 = rn * rn, r3 = rn * rn * rn;
    cplx c1 = cexp(I * K * rn) / r3;
    cplx c2 = (1.0 - I * K * rn) / r2;
    cplx c3 = K * K - 3.0 * c2;
    out[0] = c1 * (c2 * (r2 - 3.0 * xs) - K * K * (ys + zs));
    out[1] = c1 * (x * y * c3);
    out[2] = c1 * (z * x * c3);
    out[3] = c1 * (c2 * (r2 - 3.0 * ys) - K * K * (zs + xs));
    out[4] = c1 * (y * z * c3);
    out[5] = c1 * (c2 * (r2 - 3.0 * zs) - K * K * (xs + ys));
}

void orc_ldr_inter(double K, double x, double y, double z, double *out12)
{
    cplx o[6];
    ldr_inter(K, x, y, z, o);
    for (int l = 0; l < 6; l++) { out12[2 * l] = creal(o[l]); out12[2 * l + 1] = cimag(o[l]); }
}

/* ------------------------------------------------------------------------- */
/* LDR polarizability.  tools.cpp:617-630 (Get_Alpha).                         */
/* ------------------------------------------------------------------------- */
static cplx get_alpha(double K, double d, cplx diel, const double nE[3], const double nK[3])
{
    const double b1 = -1.891531, b2 = 0.1648469, b3 = -1.7700004;
    cplx a1 = (3.0 * pow(d, 3) / (4.0 * M_PI)) * (diel - 1.0) / (diel + 2.0);
    cplx res = 0.0 + (2.0 / 3.0) * pow(K * d, 3) * I;
    double S = pow(nE[0] * nK[0], 2) + pow(nE[1] * nK[1], 2) + pow(nE[2] * nK[2], 2);
    res = 1.0 + (a1 / pow(d, 3)) * ((b1 + diel * b2 + diel * b3 * S) * pow(K * d, 2) - res);
    return a1 / res;
}

void orc_get_alpha(double K, double d, double diel_re, double diel_im,
                   const double *nE, const double *nK, double *out2)
{
    cplx a = get_alpha(K, d, diel_re + I * diel_im, nE, nK);
    out2[0] = creal(a); out2[1] = cimag(a);
}

/* ------------------------------------------------------------------------- */
/* Material rule.  AProductCore.cpp:17-24: K = 2 pi / (lam / nback); each     */
/* material eps is divided by nback^2 and gets +0.01 if |eps| <= 1.01.        */
/* ------------------------------------------------------------------------- */
double orc_wavenumber(double lam, double nback) { return 2.0 * M_PI / (lam / nback); }

/* |eps| is sqrt(norm(eps)) as in the reference; strict != 0 is the periodic */
/* constructor's "< 1.01" (AProductCore.cpp:261-267).                          */
void orc_material_rule_ex(int nmat, double *mat2, double nback, int strict)
{
    for (int i = 0; i < nmat; i++) {
        cplx m = (mat2[2 * i] + I * mat2[2 * i + 1]) / (nback * nback);
        double a = sqrt(creal(m) * creal(m) + cimag(m) * cimag(m));
        if (strict ? a < 1.01 : a <= 1.01) m = m + 0.01;
        mat2[2 * i] = creal(m); mat2[2 * i + 1] = cimag(m);
    }
}

void orc_material_rule(int nmat, double *mat2, double nback) { orc_material_rule_ex(nmat, mat2, nback, 0); }

/* ------------------------------------------------------------------------- */
/* Incident field.  DDAModel.cpp:47-51 (also reset_E, DDAModel.cpp:457-468):  */
/* E[3i+c] = E0 nE[c] exp(i K d (nK . R_i)).                                  */
/* ------------------------------------------------------------------------- */
void orc_einc(int N, const int *R, double K, double d, double E0,
              const double *nK, const double *nE, double *E2)
{
    for (int i = 0; i < N; i++) {
        double ph = K * d * (nK[0] * R[3 * i] + nK[1] * R[3 * i + 1] + nK[2] * R[3 * i + 2]);
        double c = cos(ph), s = sin(ph);
        for (int k = 0; k < 3; k++) {
            E2[2 * (3 * i + k)] = E0 * nE[k] * c;
            E2[2 * (3 * i + k) + 1] = E0 * nE[k] * s;
        }
    }
}

/* ------------------------------------------------------------------------- */
/* Permittivity interpolation and inverse polarizability.                     */
/* DDAModel.cpp:54-63 (ctor), repeated in UpdateAlpha DDAModel.cpp:470-487.   */
/* diel_old holds 3N design values t; mat2 holds the rule-adjusted materials. */
/* ------------------------------------------------------------------------- */
void orc_diel_al(int n3, const double *diel_old, int nmat, const double *mat2,
                 double lam, double K, double d, const double *nE, const double *nK,
                 double *diel2, double *al2)
{
    (void)lam; (void)nmat;
    for (int i = 0; i < n3; i++) {
        int lf = (int)floor(diel_old[i]);
        int ln = lf + 1;
        if (lf >= 1) ln = lf;
        cplx mf = mat2[2 * lf] + I * mat2[2 * lf + 1];
        cplx mn = mat2[2 * ln] + I * mat2[2 * ln + 1];
        cplx e = mf + (diel_old[i] - (double)lf) * (mn - mf);
        cplx a = 1.0 / get_alpha(K, d, e, nE, nK);
        diel2[2 * i] = creal(e); diel2[2 * i + 1] = cimag(e);
        al2[2 * i] = creal(a); al2[2 * i + 1] = cimag(a);
    }
}

/* ------------------------------------------------------------------------- */
/* Local field.  DDAModel.cpp:567-575: E_int = al P / ((2 + eps) / 3).        */
/* ------------------------------------------------------------------------- */
void orc_local_field(int n3, const double *al2, const double *P2, const double *diel2, double *out2)
{
    for (int i = 0; i < n3; i++) {
        cplx lf = (2.0 + (diel2[2 * i] + I * diel2[2 * i + 1])) / 3.0;
        cplx v = (al2[2 * i] + I * al2[2 * i + 1]) * (P2[2 * i] + I * P2[2 * i + 1]) / lf;
        out2[2 * i] = creal(v); out2[2 * i + 1] = cimag(v);
    }
}

/* ------------------------------------------------------------------------- */
/* Off-diagonal product by direct summation: y_i = sum_{j != i} A(r_i - r_j) b_j */
/* This is what AProductCore::Aproduct (AProductCore.cpp:697-758) computes     */
/* through its zero-padded FFT convolution.  O(N^2); tiny lattices only.       */
/* ------------------------------------------------------------------------- */
void orc_aproduct_direct(int N, const int *R, double d, double K, const double *b2, double *y2)
{
    #pragma omp parallel for schedule(dynamic, 16)
    for (int i = 0; i < N; i++) {
        cplx acc[3] = {0, 0, 0};
        for (int j = 0; j < N; j++) {
            if (j == i) continue;
            cplx A[6];
            ldr_inter(K, d * (R[3 * i] - R[3 * j]), d * (R[3 * i + 1] - R[3 * j + 1]),
                      d * (R[3 * i + 2] - R[3 * j + 2]), A);
            cplx bx = b2[6 * j] + I * b2[6 * j + 1];
            cplx by = b2[6 * j + 2] + I * b2[6 * j + 3];
            cplx bz = b2[6 * j + 4] + I * b2[6 * j + 5];
            acc[0] += A[0] * bx + A[1] * by + A[2] * bz;
            acc[1] += A[1] * bx + A[3] * by + A[4] * bz;
            acc[2] += A[2] * bx + A[4] * by + A[5] * bz;
        }
        for (int c = 0; c < 3; c++) { y2[6 * i + 2 * c] = creal(acc[c]); y2[6 * i + 2 * c + 1] = cimag(acc[c]); }
    }
}

/* ------------------------------------------------------------------------- */
/* Mixed-radix complex FFT (Stockham, self-sorting) for the FFT-based oracle. */
/* The reference uses cuFFT Z2Z (AProductCore.cpp:47, 719-739); this is a     */
/* project-written FFT with long-double twiddles, pinned against numpy's      */
/* pocketfft and against the direct sum.                                       */
/* ------------------------------------------------------------------------- */
typedef struct {
    int L;
    int nf;
    int fac[32];
    cplx *tw; /* tw[q] = exp(-2 pi i q / L) */
} fftplan1;

static void plan1_init(fftplan1 *p, int L)
{
    p->L = L;
    p->nf = 0;
    int n = L;
    static const int pref[] = {4, 2, 3, 5, 7};
    for (int k = 0; k < 5; k++) {
        while (n % pref[k] == 0) { p->fac[p->nf++] = pref[k]; n /= pref[k]; }
    }
    for (int f = 11; n > 1; f += 2) {
        while (n % f == 0) { p->fac[p->nf++] = f; n /= f; }
    }
    p->tw = (cplx *)malloc(sizeof(cplx) * (size_t)L);
    for (int q = 0; q < L; q++) {
        long double a = -2.0L * ORC_PI * (long double)q / (long double)L;
        p->tw[q] = (double)cosl(a) + I * (double)sinl(a);
    }
}

static void plan1_free(fftplan1 *p) { free(p->tw); p->tw = NULL; }

/* In-place FFT of a contiguous pencil x[L] using scratch y[L].  sign=-1 forward,
 * sign=+1 inverse (unnormalised, as cuFFT). */
static void fft1(const fftplan1 *p, cplx *x, cplx *y, int sign)
{
    int L = p->L;
    if (L == 1) return;
    cplx *src = x, *dst = y;
    int Ns = 1;
    /* radix buffers: on the stack up to radix 64, else on the heap (the
     * reference's 2N-1 grids have prime factors up to 127 for N <= 64) */
    int rmax = 1;
    for (int s = 0; s < p->nf; s++) rmax = p->fac[s] > rmax ? p->fac[s] : rmax;
    cplx vs[64], ws[64];
    cplx *v = vs, *w = ws;
    if (rmax > 64) {
        v = (cplx *)malloc(sizeof(cplx) * (size_t)rmax);
        w = (cplx *)malloc(sizeof(cplx) * (size_t)rmax);
    }
    for (int s = 0; s < p->nf; s++) {
        int R = p->fac[s];
        int LR = L / R;
        int tws = L / (Ns * R);
        for (int j = 0; j < LR; j++) {
            int k = j % Ns;
            for (int m = 0; m < R; m++) {
                cplx t = p->tw[(int)(((int64_t)m * k * tws) % L)];
                if (sign > 0) t = conj(t);
                v[m] = src[j + m * LR] * t;
            }
            for (int q = 0; q < R; q++) {
                cplx acc = 0;
                for (int m = 0; m < R; m++) {
                    cplx t = p->tw[(int)(((int64_t)m * q % R) * LR)];
                    if (sign > 0) t = conj(t);
                    acc += v[m] * t;
                }
                w[q] = acc;
            }
            int base = (j / Ns) * Ns * R + k;
            for (int q = 0; q < R; q++) dst[base + q * Ns] = w[q];
        }
        cplx *tmp = src; src = dst; dst = tmp;
        Ns *= R;
    }
    if (v != vs) {
        free(v);
        free(w);
    }
    if (src != x) memcpy(x, src, sizeof(cplx) * (size_t)L);
}

void orc_fft1d(int L, double *x2, int sign)
{
    fftplan1 p;
    plan1_init(&p, L);
    cplx *y = (cplx *)malloc(sizeof(cplx) * (size_t)L);
    fft1(&p, (cplx *)x2, y, sign);
    free(y);
    plan1_free(&p);
}

/* ------------------------------------------------------------------------- */
/* FFT-based operator.  Restates AProductCore.cpp:12-249 (grid, Green fill,   */
/* spectrum) and AProductCore.cpp:697-758 (scatter, FFT, 6x3 symmetric        */
/* product of kernel.cu:74-95, inverse FFT, gather, divide by grid size).     */
/* Grid (Mx, My, Mz) may be any size >= (2Nx-1, 2Ny-1, 2Nz-1); the reference  */
/* uses exactly 2N-1 (AProductCore.cpp:41-44).                                */
/* ------------------------------------------------------------------------- */
typedef struct {
    int N, Nx, Ny, Nz, Mx, My, Mz;
    size_t M;
    int *R;
    cplx *G[6];     /* Green spectrum, full grid, [x][y][z] z fastest */
    cplx *B[3];     /* work grids */
    fftplan1 px, py, pz;
} orc_op;

static int grid_offset(int i, int Nn, int Mm, int *ok)
{
    /* AProductCore.cpp:54-194: index i < Nn -> offset i; i >= Mm-(Nn-1) -> i-Mm. */
    *ok = 1;
    if (i <= Nn - 1) return i;
    if (i >= Mm - (Nn - 1)) return i - Mm;
    *ok = 0;
    return 0;
}

static void fft3_axis(orc_op *op, cplx *g, int axis, int sign, int x_lim, int y_lim)
{
    /* transforms along one axis; pencils restricted to x < x_lim, y < y_lim where
     * the pruned convolution leaves the rest zero / unused */
    int Mx = op->Mx, My = op->My, Mz = op->Mz;
    if (axis == 2) {
        #pragma omp parallel
        {
            cplx *s = (cplx *)malloc(sizeof(cplx) * (size_t)Mz);
            #pragma omp for collapse(2) schedule(static)
            for (int x = 0; x < x_lim; x++)
                for (int y = 0; y < y_lim; y++)
                    fft1(&op->pz, g + ((size_t)x * My + y) * Mz, s, sign);
            free(s);
        }
    } else if (axis == 1) {
        #pragma omp parallel
        {
            cplx *a = (cplx *)malloc(sizeof(cplx) * (size_t)My);
            cplx *s = (cplx *)malloc(sizeof(cplx) * (size_t)My);
            #pragma omp for collapse(2) schedule(static)
            for (int x = 0; x < x_lim; x++)
                for (int z = 0; z < Mz; z++) {
                    cplx *base = g + (size_t)x * My * Mz + z;
                    for (int y = 0; y < My; y++) a[y] = base[(size_t)y * Mz];
                    fft1(&op->py, a, s, sign);
                    for (int y = 0; y < My; y++) base[(size_t)y * Mz] = a[y];
                }
            free(a); free(s);
        }
    } else {
        #pragma omp parallel
        {
            cplx *a = (cplx *)malloc(sizeof(cplx) * (size_t)Mx);
            cplx *s = (cplx *)malloc(sizeof(cplx) * (size_t)Mx);
            #pragma omp for collapse(2) schedule(static)
            for (int y = 0; y < My; y++)
                for (int z = 0; z < Mz; z++) {
                    cplx *base = g + (size_t)y * Mz + z;
                    for (int x = 0; x < Mx; x++) a[x] = base[(size_t)x * My * Mz];
                    fft1(&op->px, a, s, sign);
                    for (int x = 0; x < Mx; x++) base[(size_t)x * My * Mz] = a[x];
                }
            free(a); free(s);
        }
    }
}

void *orc_op_create(int N, const int *R, int Nx, int Ny, int Nz, int Mx, int My, int Mz,
                    double d, double K)
{
    if (Mx < 2 * Nx - 1 || My < 2 * Ny - 1 || Mz < 2 * Nz - 1) return NULL;
    orc_op *op = (orc_op *)calloc(1, sizeof(orc_op));
    op->N = N; op->Nx = Nx; op->Ny = Ny; op->Nz = Nz;
    op->Mx = Mx; op->My = My; op->Mz = Mz;
    op->M = (size_t)Mx * My * Mz;
    op->R = (int *)malloc(sizeof(int) * 3 * (size_t)N);
    memcpy(op->R, R, sizeof(int) * 3 * (size_t)N);
    plan1_init(&op->px, Mx); plan1_init(&op->py, My); plan1_init(&op->pz, Mz);
    for (int l = 0; l < 6; l++) op->G[l] = (cplx *)calloc(op->M, sizeof(cplx));
    for (int c = 0; c < 3; c++) op->B[c] = (cplx *)calloc(op->M, sizeof(cplx));
    #pragma omp parallel for collapse(2) schedule(static)
    for (int i = 0; i < Mx; i++)
        for (int j = 0; j < My; j++) {
            int oki, okj, okk;
            int oi = grid_offset(i, Nx, Mx, &oki);
            int oj = grid_offset(j, Ny, My, &okj);
            for (int k = 0; k < Mz; k++) {
                int ok_ = grid_offset(k, Nz, Mz, &okk);
                if (!(oki && okj && okk)) continue;
                cplx A[6];
                ldr_inter(K, d * oi, d * oj, d * ok_, A);
                size_t idx = ((size_t)i * My + j) * Mz + k;
                for (int l = 0; l < 6; l++) op->G[l][idx] = A[l];
            }
        }
    for (int l = 0; l < 6; l++) {
        fft3_axis(op, op->G[l], 2, -1, Mx, My);
        fft3_axis(op, op->G[l], 1, -1, Mx, My);
        fft3_axis(op, op->G[l], 0, -1, Mx, My);
    }
    return op;
}

void orc_op_destroy(void *h)
{
    orc_op *op = (orc_op *)h;
    if (!op) return;
    for (int l = 0; l < 6; l++) free(op->G[l]);
    for (int c = 0; c < 3; c++) free(op->B[c]);
    plan1_free(&op->px); plan1_free(&op->py); plan1_free(&op->pz);
    free(op->R);
    free(op);
}

/* y = A_offdiag b  (AProductCore.cpp:697-758); if al2 != NULL also adds al o b
 * (DDAModel::Aproductwithalb, DDAModel.cpp:577-586). */
static void op_apply(orc_op *op, const cplx *b, const cplx *al, cplx *y)
{
    size_t M = op->M;
    int Mx = op->Mx, My = op->My, Mz = op->Mz;
    for (int c = 0; c < 3; c++) memset(op->B[c], 0, sizeof(cplx) * M);
    for (int i = 0; i < op->N; i++) {
        size_t idx = ((size_t)op->R[3 * i] * My + op->R[3 * i + 1]) * Mz + op->R[3 * i + 2];
        for (int c = 0; c < 3; c++) op->B[c][idx] = b[3 * i + c];
    }
    for (int c = 0; c < 3; c++) {
        fft3_axis(op, op->B[c], 2, -1, op->Nx, op->Ny);
        fft3_axis(op, op->B[c], 1, -1, op->Nx, My);
        fft3_axis(op, op->B[c], 0, -1, Mx, My);
    }
    /* kernel.cu:74-95 symmetric product */
    #pragma omp parallel for schedule(static)
    for (size_t q = 0; q < M; q++) {
        cplx bx = op->B[0][q], by = op->B[1][q], bz = op->B[2][q];
        op->B[0][q] = op->G[0][q] * bx + op->G[1][q] * by + op->G[2][q] * bz;
        op->B[1][q] = op->G[1][q] * bx + op->G[3][q] * by + op->G[4][q] * bz;
        op->B[2][q] = op->G[2][q] * bx + op->G[4][q] * by + op->G[5][q] * bz;
    }
    for (int c = 0; c < 3; c++) {
        fft3_axis(op, op->B[c], 0, +1, Mx, My);
        fft3_axis(op, op->B[c], 1, +1, op->Nx, My);
        fft3_axis(op, op->B[c], 2, +1, op->Nx, op->Ny);
    }
    double inv = 1.0 / (double)M;
    #pragma omp parallel for schedule(static)
    for (int i = 0; i < op->N; i++) {
        size_t idx = ((size_t)op->R[3 * i] * My + op->R[3 * i + 1]) * Mz + op->R[3 * i + 2];
        for (int c = 0; c < 3; c++) {
            cplx v = op->B[c][idx] / (double)M;
            (void)inv;
            if (al) v += al[3 * i + c] * b[3 * i + c];
            y[3 * i + c] = v;
        }
    }
}

void orc_op_apply(void *h, const double *b2, const double *al2, double *y2)
{
    op_apply((orc_op *)h, (const cplx *)b2, (const cplx *)al2, (cplx *)y2);
}

/* ------------------------------------------------------------------------- */
/* Sesquilinear dot, Eigen's a.dot(b) = sum conj(a_i) b_i                     */
/* (eigen-eigen-94875feeeeb9/Eigen/src/Core/Dot.h:61-72), and norm.           */
/* ------------------------------------------------------------------------- */
static cplx vdot(size_t n, const cplx *a, const cplx *b)
{
    cplx s = 0;
    for (size_t i = 0; i < n; i++) s += conj(a[i]) * b[i];
    return s;
}

static double vnorm(size_t n, const cplx *a)
{
    double s = 0;
    for (size_t i = 0; i < n; i++) s += creal(a[i]) * creal(a[i]) + cimag(a[i]) * cimag(a[i]);
    return sqrt(s);
}

/* ------------------------------------------------------------------------- */
/* BiCGSTAB.  DDAModel.cpp:131-260, statement order kept: unrolled first step */
/* (:164-182), loop (:186-255) with the r0.rl == 0 breakdown test and the      */
/* recurrence-residual stop ||r||/||E|| <= tol.  P (x0 in, solution out).     */
/* Returns 0 converged, 1 breakdown, 2 not converged.  iters = ITERATION.     */
/* max_aproducts > 0 stops early after that many operator applications       */
/* (used to time a bounded sample); then the return value is 3.              */
/* ------------------------------------------------------------------------- */
int orc_bicgstab(void *h, const double *al2, const double *E2, double *P2,
                 int max_it, double tol, int max_aproducts, int *iters, double *err)
{
    orc_op *op = (orc_op *)h;
    size_t n = 3 * (size_t)op->N;
    const cplx *al = (const cplx *)al2, *E = (const cplx *)E2;
    cplx *P = (cplx *)P2;
    cplx *p = calloc(n, sizeof(cplx)), *t = calloc(n, sizeof(cplx)), *r = calloc(n, sizeof(cplx));
    cplx *r0 = calloc(n, sizeof(cplx)), *rl = calloc(n, sizeof(cplx)), *u = calloc(n, sizeof(cplx));
    cplx *Ap = calloc(n, sizeof(cplx)), *At = calloc(n, sizeof(cplx));
    int status = 2, napp = 0;
    double En = vnorm(n, E);
    cplx alpha, beta, zeta;
    *iters = 0;
    *err = 0;

    op_apply(op, P, al, Ap); napp++;                 /* Ax0 */
    for (size_t i = 0; i < n; i++) { r[i] = E[i] - Ap[i]; r0[i] = r[i]; p[i] = r[i]; }
    op_apply(op, p, al, Ap); napp++;                 /* Ap0 */
    alpha = vdot(n, r0, r) / vdot(n, r0, Ap);
    for (size_t i = 0; i < n; i++) t[i] = r[i] - alpha * Ap[i];
    op_apply(op, t, al, At); napp++;                 /* At0 */
    zeta = vdot(n, At, t) / vdot(n, At, At);
    for (size_t i = 0; i < n; i++) {
        u[i] = zeta * Ap[i];
        cplx z = zeta * r[i] - alpha * u[i];
        P[i] = P[i] + alpha * p[i] + z;
        rl[i] = r[i];
        r[i] = t[i] - zeta * At[i];
    }
    for (int it = 0; it < max_it; it++) {
        if (max_aproducts > 0 && napp + 2 > max_aproducts) { status = 3; *iters = it; break; }
        cplx r0rl = vdot(n, r0, rl);
        if (creal(r0rl) == 0 && cimag(r0rl) == 0) {
            *err = vnorm(n, r) / En;
            *iters = it + 1;
            status = 1;
            break;
        }
        beta = (alpha / zeta) * vdot(n, r0, r) / r0rl;
        for (size_t i = 0; i < n; i++) p[i] = r[i] + beta * (p[i] - u[i]);
        op_apply(op, p, al, Ap); napp++;
        alpha = vdot(n, r0, r) / vdot(n, r0, Ap);
        for (size_t i = 0; i < n; i++) t[i] = r[i] - alpha * Ap[i];
        op_apply(op, t, al, At); napp++;
        zeta = vdot(n, At, t) / vdot(n, At, At);
        for (size_t i = 0; i < n; i++) {
            u[i] = zeta * Ap[i];
            cplx z = zeta * r[i] - alpha * u[i];
            P[i] = P[i] + alpha * p[i] + z;
            rl[i] = r[i];
            r[i] = t[i] - zeta * At[i];
        }
        double e = vnorm(n, r) / En;
        *err = e;
        *iters = it + 1;
        if (e <= tol) { status = 0; break; }
    }
    free(p); free(t); free(r); free(r0); free(rl); free(u); free(Ap); free(At);
    return status;
}

void orc_set_num_threads(int n)
{
    if (n > 0) omp_set_num_threads(n);
}

int orc_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
