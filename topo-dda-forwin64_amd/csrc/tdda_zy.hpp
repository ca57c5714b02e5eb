// tdda_zy.hpp -- fused z + y passes for slabs that fit in LDS.
//
// The separate passes move every x slab through HBM twice between the z and
// the y transforms: P1 writes w1[x][y][kz][c] (Nx Ny Mz 3 elements) and P2
// reads it back, P4 writes it again and P5 reads it.  For the small
// structures of a batch (Ny <= 64, Mz <= 32) a whole x slab after the z
// transform fits in LDS, so one workgroup per slab does both transforms:
//
//   k_pzy (P1 + P2): site vector slab (+ the BiCGSTAB vector update of P1)
//                    -> z transform in registers -> slab in LDS
//                    -> y transform (register four-step)  -> w2[x][ky][kz][c]
//   k_pyz (P4 + P5): w2[x][ky][kz][c] -> inverse y transform -> slab in LDS
//                    -> inverse z transform in registers -> crop, diagonal,
//                       solver epilogue and reductions of P5
//
// w1 never reaches HBM: per application 4 Nx Ny Mz 3 elements less traffic
// and two launches less.  The x pass is unchanged.
//
// The z transform of a pencil (y, c) runs in one thread: all Mz points in
// registers (sdft, compile-time twiddles), selected per job by a uniform
// switch over the compiled z lengths, so the kernels are specialised on the
// y length only.  The y transform is the four-step split of k_pyr with its
// first (forward) or last (inverse) step reading or writing the LDS slab.
//
// Tile = one x slab of one job (pz = Ny pencils in the job descriptor, so
// the reducing pass keeps one partial-sum slot per slab).
#pragma once
#include "tdda_kernels.hpp"
#include "tdda_zy_plan.hpp"

namespace tdda {

// Forward z transform of one pencil: src[3 z] for z < Nz (zero above),
// dst[3 kz] for kz < LZ.
template <typename T, int LZ>
__device__ __forceinline__ void zy_zfwd(const Cx<T>* src, Cx<T>* dst, int Nz)
{
    constexpr int H = (LZ + 1) / 2;  // Nz <= (Mz + 1) / 2: the upper half is zero padding
    Cx<T> v[LZ];
    sfor<0, LZ>([&](auto nc) {
        constexpr int n = decltype(nc)::value;
        if constexpr (n >= H) v[n] = Cx<T>{T(0), T(0)};
        else v[n] = n < Nz ? src[3 * n] : Cx<T>{T(0), T(0)};
    });
    sdft<T, LZ, false>(v);
    sfor<0, LZ>([&](auto kc) {
        constexpr int k = decltype(kc)::value;
        dst[3 * k] = v[k];
    });
}

// Inverse z transform of one pencil: src[3 kz] for kz < LZ, dst[3 z] for
// z < Nz.
template <typename T, int LZ>
__device__ __forceinline__ void zy_zinv(const Cx<T>* src, Cx<T>* dst, int Nz)
{
    constexpr int H = (LZ + 1) / 2;
    Cx<T> v[LZ];
    sfor<0, LZ>([&](auto kc) {
        constexpr int k = decltype(kc)::value;
        v[k] = src[3 * k];
    });
    sdft<T, LZ, true>(v);
    sfor<0, H>([&](auto nc) {
        constexpr int n = decltype(nc)::value;
        if (n < Nz) dst[3 * n] = v[n];
    });
}

// Site vector index of element e of the slab whose first cell is `base`
// (e = (y Nz + z) 3 + c), or -1 for an empty cell.
template <typename T>
__device__ __forceinline__ int zy_index(const JobDesc<T>* J, int base, int e)
{
    if (!J->cellmap) return base * 3 + e;
    const int site = gldi(J->cellmap + base + e / 3);
    return site >= 0 ? site * 3 + (e - (e / 3) * 3) : -1;
}

// Elements per thread whose global loads the staging and epilogue loops
// issue together (one memory round trip per group instead of per element);
// complex128 takes two, within its registers.
template <typename T>
__host__ __device__ constexpr int zy_group()
{
    return sizeof(T) == 8 ? 2 : 4;
}

template <typename T, bool INV>
__device__ __forceinline__ void zy_zpencil(int Mz, const Cx<T>* src, Cx<T>* dst, int Nz)
{
    switch (Mz) {
#define TDDA_CASE(LL)                                          \
    case LL:                                                   \
        if constexpr (LL <= zy_zmax<T>()) {                    \
            if constexpr (INV) zy_zinv<T, LL>(src, dst, Nz);   \
            else zy_zfwd<T, LL>(src, dst, Nz);                 \
        }                                                      \
        break;
        TDDA_ZY_ZLENGTHS(TDDA_CASE)
#undef TDDA_CASE
    default: break;  // the planner fuses only the compiled z lengths
    }
}

// (job, x slab) of fused tile tl by the tile starts.
__device__ __forceinline__ int2 zy_tile(const int* __restrict__ tstart, int njobs, int tl)
{
    const int job = find_job(tstart, njobs, tl);
    return make_int2(job, tl - tstart[job]);
}

// P1 + P2 fused.  mode: IN_PLAIN, IN_X, IN_P, IN_T (as k_p1).
template <typename T, int LY>
__global__ void __launch_bounds__(kThreads, (zy_waves<T, false>())) k_pzy(const JobDesc<T>* __restrict__ jobs,
                                                                const int* __restrict__ tstart, int njobs, int mode,
                                                                const int2* __restrict__ tmap)
{
    constexpr int L1 = pxr_l1(LY), L2 = LY / L1, H2 = (L2 + 1) / 2;
    static_assert(L1 > 0 && L1 >= L2, "length without a register split");
    extern __shared__ __align__(16) char smem[];
    Cx<T>* tw = reinterpret_cast<Cx<T>*>(smem);  // tw[q] = exp(-2 pi i q / LY)
    Cx<T>* ZB = tw + pass_tw_slots(true, LY, 0);
    const int ntot = tstart[njobs];
    for (int i = threadIdx.x; i < LY; i += kThreads) tw[i] = gld(jobs[0].twfy + i);
    for (int tl = blockIdx.x; tl < ntot; tl += gridDim.x) {
        // (job, x slab) of the tile: the host's tile map, else the search
        const int2 jx = tmap ? tmap[tl] : zy_tile(tstart, njobs, tl);
        const int job = jx.x;
        const JobDesc<T>* J = jobs + job;
        if (mode != IN_PLAIN && job_inactive(J->st)) continue;
        const int x = jx.y;
        const int Ny = J->Ny, Nz = J->Nz, Mz = J->Mz, F = 3 * Mz, rs = zy_row_stride(Mz);
        Cx<T>* S = ZB + Ny * rs;
        // 1. the slab's inputs -> S[y][z][c], with the fused vector updates
        Cx<T> sa{T(0), T(0)}, sb{T(0), T(0)}, sc{T(0), T(0)};
        bool first = false;
        if (mode == IN_P) {
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
        } else if (mode == IN_T) {
            sa = from_d<T>(gstate(J->st)->alpha[0], gstate(J->st)->alpha[1]);
        }
        const int nin = Ny * Nz * 3;
        const int base = x * Ny * Nz;  // cell index of (x, 0, 0)
        // loads: a = in / x / r / t; b = Ap (IN_T) or At (IN_P); IN_P after the
        // first step also c = x, d = p, f = Ap (the pending update, then p);
        // with right Jacobi preconditioning g = al (the operator input is
        // p / al or t / al)
        const bool upd = mode == IN_P && !first;
        const bool prec = (mode == IN_P || mode == IN_T) && gstate(J->st)->prec != 0;
        const Cx<T>* va = mode == IN_PLAIN ? J->in : mode == IN_X ? J->x : upd ? J->t : J->r;
        const Cx<T>* vb = mode == IN_T ? J->ap : J->at;
        const bool two = mode == IN_T || upd;
        for (int e0 = threadIdx.x; e0 < nin; e0 += zy_group<T>() * kThreads) {
            int idx[zy_group<T>()];
            Cx<T> a[zy_group<T>()], b[zy_group<T>()], c[zy_group<T>()], d[zy_group<T>()], f[zy_group<T>()],
                g[zy_group<T>()];
#pragma unroll
            for (int u = 0; u < zy_group<T>(); u++) {
                const int e = e0 + u * kThreads;
                idx[u] = e < nin ? zy_index(J, base, e) : -1;
                a[u] = b[u] = c[u] = d[u] = f[u] = g[u] = Cx<T>{T(0), T(0)};
                if (idx[u] >= 0) {
                    a[u] = gld(va + idx[u]);
                    if (two) b[u] = gld(vb + idx[u]);
                    if (prec) g[u] = gld(J->al + idx[u]);
                    if (upd) {
                        c[u] = gld(J->x + idx[u]);
                        d[u] = gld(J->p + idx[u]);
                        f[u] = gld(J->ap + idx[u]);
                    }
                }
            }
#pragma unroll
            for (int u = 0; u < zy_group<T>(); u++) {
                const int e = e0 + u * kThreads;
                if (e >= nin) break;
                Cx<T> v = a[u];
                if (idx[u] >= 0) {
                    if (mode == IN_P) {
                        if (upd) {
                            // the previous iteration's P += alpha p + z (z = zeta r - alpha u =
                            // zeta t) and r = t - zeta At (DDAModel.cpp:226-230), then
                            // p = r + beta (p - u), u = zeta Ap   (DDAModel.cpp:219)
                            const Cx<T> rn = csub(a[u], cmul(sb, b[u]));
                            gst(J->r + idx[u], rn);
                            gst(J->x + idx[u], cadd(c[u], cadd(cmul(sc, d[u]), cmul(sb, a[u]))));
                            v = cadd(rn, cmul(sa, csub(d[u], cmul(sb, f[u]))));
                        }
                        gst(J->p + idx[u], v);
                    } else if (mode == IN_T) {
                        // t = r - alpha Ap   (DDAModel.cpp:222)
                        v = csub(v, cmul(sa, b[u]));
                        gst(J->t + idx[u], v);
                    }
                    if (prec) v = cdivv(v, g[u]);
                }
                S[e] = v;
            }
        }
        __syncthreads();
        // 2. z transform of the pencils (y, c) -> ZB[y][kz][c]
        for (int pc = threadIdx.x; pc < Ny * 3; pc += kThreads) {
            const int y = pc / 3, c = pc - 3 * y;
            zy_zpencil<T, false>(Mz, S + y * Nz * 3 + c, ZB + y * rs + c, Nz);
        }
        __syncthreads();
        // 3. y transform of the slab's F columns, nf at a time (k_pyr's
        // forward four-step with the first step reading the slab)
        const int nf = J->tfy, ps = J->ps_y;
        Cx<T>* A = S;  // the staged inputs are no longer needed
        Cx<T>* w2 = J->w2 + (size_t)x * LY * F;
        const int q = threadIdx.x / nf, col = threadIdx.x - q * nf;
        for (int f0 = 0; f0 < F; f0 += nf) {
            const bool on = f0 + col < F;
            if (q < L1 && on) {
                const int n1 = q;
                Cx<T> v[L2];
                sfor<0, L2>([&](auto c2) {
                    constexpr int n2 = decltype(c2)::value;
                    if constexpr (n2 >= H2) {
                        v[n2] = Cx<T>{T(0), T(0)};
                    } else {
                        const int y = n1 + L1 * n2;
                        v[n2] = y < Ny ? ZB[y * rs + f0 + col] : Cx<T>{T(0), T(0)};
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
            if (q < L2 && on) {
                const int k2 = q;
                Cx<T> v[L1];
                const Cx<T>* cp = A + col * ps + k2;
                sfor<0, L1>([&](auto c1) { v[decltype(c1)::value] = cp[decltype(c1)::value * L2]; });
                sdft<T, L1, false>(v);
                sfor<0, L1>([&](auto c1) {
                    constexpr int k1 = decltype(c1)::value;
                    gst(w2 + ((k2 + L2 * k1) * F + f0 + col), v[k1]);
                });
            }
            __syncthreads();
        }
    }
}

// P4 + P5 fused.  mode: OUT_PLAIN (diag: + al o in), OUT_RES, OUT_AP,
// OUT_AT (as k_p5).
template <typename T, int LY>
__global__ void __launch_bounds__(kThreads, (zy_waves<T, true>())) k_pyz(const JobDesc<T>* __restrict__ jobs,
                                                                const int* __restrict__ tstart, int njobs, int mode,
                                                                int diag, const int2* __restrict__ tmap)
{
    constexpr int L1 = pxr_l1(LY), L2 = LY / L1, H2 = (L2 + 1) / 2;
    static_assert(L1 > 0 && L1 >= L2, "length without a register split");
    extern __shared__ __align__(16) char smem[];
    __shared__ double red[16 * kPart];
    Cx<T>* tw = reinterpret_cast<Cx<T>*>(smem);
    Cx<T>* ZB = tw + pass_tw_slots(true, LY, 0);
    const int ntot = tstart[njobs];
    for (int i = threadIdx.x; i < LY; i += kThreads) tw[i] = gld(jobs[0].twfy + i);
    __syncthreads();
    for (int tl = blockIdx.x; tl < ntot; tl += gridDim.x) {
        const int2 jx = tmap ? tmap[tl] : zy_tile(tstart, njobs, tl);
        const int job = jx.x;
        const JobDesc<T>* J = jobs + job;
        if (mode != OUT_PLAIN && job_inactive(J->st)) continue;
        const int x = jx.y;
        const int Ny = J->Ny, Nz = J->Nz, Mz = J->Mz, F = 3 * Mz, rs = zy_row_stride(Mz);
        Cx<T>* S = ZB + Ny * rs;
        Cx<T>* A = S;
        // 1. inverse y transform of the F columns, nf at a time -> ZB[y < Ny]
        const int nf = J->tfy, ps = J->ps_y;
        const Cx<T>* w2 = J->w2 + (size_t)x * LY * F;
        const int q = threadIdx.x / nf, col = threadIdx.x - q * nf;
        // complex64: the loads of the next column block are issued before
        // the current block is transformed (complex128 has no registers to
        // spare for them)
        constexpr bool PF = sizeof(T) == 4;
        Cx<T> nx[L1];
        auto load_block = [&](int f0, Cx<T>* dst) {
            if (q < L2 && f0 + col < F) {
                sfor<0, L1>([&](auto c1) {
                    constexpr int k1 = decltype(c1)::value;
                    dst[k1] = gld(w2 + ((q + L2 * k1) * F + f0 + col));
                });
            }
        };
        if (PF) load_block(0, nx);
        for (int f0 = 0; f0 < F; f0 += nf) {
            const bool on = f0 + col < F;
            Cx<T> v[L1];
            if constexpr (PF) {
                sfor<0, L1>([&](auto c1) { v[decltype(c1)::value] = nx[decltype(c1)::value]; });
                if (f0 + nf < F) load_block(f0 + nf, nx);
            } else {
                load_block(f0, v);
            }
            if (q < L2 && on) {
                const int k2 = q;
                sdft<T, L1, true>(v);
                Cx<T>* cp = A + col * ps + k2;
                cp[0] = v[0];
                sfor<1, L1>([&](auto c1) {
                    constexpr int n1 = decltype(c1)::value;
                    cp[n1 * L2] = cmulc(v[n1], tw[n1 * k2]);
                });
            }
            __syncthreads();
            if (q < L1 && on) {
                const int n1 = q;
                Cx<T> v[L2];
                const Cx<T>* row = A + col * ps + n1 * L2;
                sfor<0, L2>([&](auto c2) { v[decltype(c2)::value] = row[decltype(c2)::value]; });
                sdft<T, L2, true>(v);
                sfor<0, H2>([&](auto c2) {
                    constexpr int n2 = decltype(c2)::value;
                    const int y = n1 + L1 * n2;
                    if (y < Ny) ZB[y * rs + f0 + col] = v[n2];
                });
            }
            __syncthreads();
        }
        // 2. inverse z transform of the pencils (y, c) -> S[y][z][c], z < Nz
        for (int pc = threadIdx.x; pc < Ny * 3; pc += kThreads) {
            const int y = pc / 3, c = pc - 3 * y;
            zy_zpencil<T, true>(Mz, ZB + y * rs + c, S + y * Nz * 3 + c, Nz);
        }
        __syncthreads();
        // 3. crop to the sites, diagonal and solver epilogue (k_p5's)
        double acc[kPart] = {};
        const int nout = Ny * Nz * 3;
        const int base = x * Ny * Nz;
        // per mode: a = al, b = the vector al multiplies, c = E (OUT_RES) or
        // r0 (OUT_AP, OUT_AT)
        // right Jacobi preconditioning: the operator is A D^-1, whose
        // diagonal term on v is v itself (no al)
        const bool prec = (mode == OUT_AP || mode == OUT_AT) && gstate(J->st)->prec != 0;
        const bool use_al = (mode != OUT_PLAIN || diag) && !prec;
        const Cx<T>* vb = mode == OUT_PLAIN ? J->in : mode == OUT_RES ? J->x : mode == OUT_AP ? J->p : J->t;
        const Cx<T>* vc = mode == OUT_RES ? J->e : J->r0;
        const bool use_c = mode != OUT_PLAIN;
        for (int e0 = threadIdx.x; e0 < nout; e0 += zy_group<T>() * kThreads) {
            int idx[zy_group<T>()];
            Cx<T> a[zy_group<T>()], b[zy_group<T>()], c[zy_group<T>()];
#pragma unroll
            for (int u = 0; u < zy_group<T>(); u++) {
                const int e = e0 + u * kThreads;
                idx[u] = e < nout ? zy_index(J, base, e) : -1;
                a[u] = b[u] = c[u] = Cx<T>{T(0), T(0)};
                if (idx[u] >= 0) {
                    if (use_al) a[u] = gld(J->al + idx[u]);
                    if (use_al || prec) b[u] = gld(vb + idx[u]);
                    if (use_c) c[u] = gld(vc + idx[u]);
                }
            }
#pragma unroll
            for (int u = 0; u < zy_group<T>(); u++) {
                const int i = idx[u];
                if (i < 0) continue;
                const Cx<T> conv = S[e0 + u * kThreads];
                if (mode == OUT_PLAIN) {
                    gst(J->out + i, use_al ? cadd(conv, cmul(a[u], b[u])) : conv);
                } else if (mode == OUT_RES) {
                    // r = E - A x ; r0 = r   (DDAModel.cpp:163-166)
                    const Cx<T> r = csub(c[u], cadd(conv, cmul(a[u], b[u])));
                    gst(J->r + i, r);
                    gst(J->r0 + i, r);
                    const double2 rd = to_d(r), ed = to_d(c[u]);
                    acc[0] += rd.x * rd.x + rd.y * rd.y;
                    acc[1] += ed.x * ed.x + ed.y * ed.y;
                } else if (mode == OUT_AP) {
                    const Cx<T> ap = cadd(conv, prec ? b[u] : cmul(a[u], b[u]));
                    gst(J->ap + i, ap);
                    const double2 r0 = to_d(c[u]), bd = to_d(ap);
                    acc[0] += r0.x * bd.x + r0.y * bd.y;  // conj(r0) Ap
                    acc[1] += r0.x * bd.y - r0.y * bd.x;
                } else {  // OUT_AT
                    const Cx<T> at = cadd(conv, prec ? b[u] : cmul(a[u], b[u]));
                    gst(J->at + i, at);
                    at_sums(acc, at, b[u], c[u]);
                }
            }
        }
        if (mode != OUT_PLAIN) {
            block_sum<kPart>(acc, red);
            store_partials(J->part, x, acc);
        }
        __syncthreads();
    }
}

}  // namespace tdda
