// tdda_capi.hip -- C ABI of the DDA solve engine: contexts, operators (Green
// spectrum on the device), models (device-resident solver state) and the
// solver drivers.  Declarations and reference mapping: include/topodda.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <complex>
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <type_traits>
#include <vector>

#include "topodda.h"
#include "tdda_kernels.hpp"
#include "tdda_setup.hpp"
#include "tdda_launch.hpp"
#include "tdda_zy_plan.hpp"
#include "tdda_grid.hpp"

using namespace tdda;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg)
{
    g_err = msg;
    return code;
}

}  // namespace

// The library's error message for its other translation units.
namespace tdda {
int set_error(int code, const char* msg) { return fail(code, msg); }
}  // namespace tdda

namespace {

#define HIPCHK(x)                                                                                  \
    do {                                                                                           \
        hipError_t e_ = (x);                                                                       \
        if (e_ != hipSuccess) return fail(TDDA_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

#define CHK(x)                 \
    do {                       \
        int rc_ = (x);         \
        if (rc_ != TDDA_OK) return rc_; \
    } while (0)

bool smooth235(int n)
{
    if (n <= 0) return false;
    for (int p : {2, 3, 5})
        while (n % p == 0) n /= p;
    return n == 1;
}

int next_smooth(int n)
{
    if (n < 1) n = 1;
    while (!smooth235(n)) n++;
    return n;
}

FftPlan make_plan(int L)
{
    FftPlan p{};
    p.L = L;
    int n = L, e2 = 0;
    while (n % 2 == 0) {
        n /= 2;
        e2++;
    }
    while (e2 >= 3) {
        p.rad[p.nst++] = 8;
        e2 -= 3;
    }
    if (e2 == 2) p.rad[p.nst++] = 4;
    if (e2 == 1) p.rad[p.nst++] = 2;
    while (n % 3 == 0) {
        p.rad[p.nst++] = 3;
        n /= 3;
    }
    while (n % 5 == 0) {
        p.rad[p.nst++] = 5;
        n /= 5;
    }
    p.tb = 1;
    while (p.tb * p.tb < L) p.tb++;
    return p;
}

// two-level table: T0[q0] = w^q0 (q0 < tb), T1[q1] = w^(q1 tb) (q1 < ceil(L / tb))
std::vector<Cx<double>> twiddles2(int L, int tb)
{
    const int n1 = (L + tb - 1) / tb;
    std::vector<Cx<double>> t(tb + n1);
    const long double pi = 3.141592653589793238462643383279502884L;
    for (int q = 0; q < tb; q++) {
        const long double a = -2.0L * pi * (long double)q / (long double)L;
        t[q] = Cx<double>{(double)cosl(a), (double)sinl(a)};
    }
    for (int q = 0; q < n1; q++) {
        const long double a = -2.0L * pi * (long double)(q * tb) / (long double)L;
        t[tb + q] = Cx<double>{(double)cosl(a), (double)sinl(a)};
    }
    return t;
}

// roots of unity exp(-2 pi i q / L) in long double
std::vector<Cx<double>> twiddles(int L)
{
    std::vector<Cx<double>> t(L);
    const long double pi = 3.141592653589793238462643383279502884L;
    for (int q = 0; q < L; q++) {
        const long double a = -2.0L * pi * (long double)q / (long double)L;
        t[q] = Cx<double>{(double)cosl(a), (double)sinl(a)};
    }
    return t;
}

int env_int(const char* name, int dflt)
{
    const char* v = getenv(name);
    return v ? atoi(v) : dflt;
}

}  // namespace

// ---------------------------------------------------------------------------
// handles
// ---------------------------------------------------------------------------
// Handles are reference counted so that destroying them in any order is
// safe: a model keeps its operator alive, an operator keeps its context.
// Streams a pass fans its launch groups out over (besides the context's
// own), so that the launches of different FFT lengths run concurrently and
// fill each other's tails; fork and join are events on the main stream.
constexpr int kAuxStreams = 3;
struct AuxStreams {
    int n = 0;
    hipStream_t s[kAuxStreams] = {};
    hipEvent_t fork = nullptr;
    hipEvent_t join[kAuxStreams] = {};
};

struct tdda_ctx_s {
    int device;
    hipStream_t stream;
    size_t lds_budget;  // per workgroup
    size_t lds_cu;      // per CU
    int ncu;
    int refs;
    // stream-ordered allocation (request loop): hipFreeAsync does not
    // synchronise the device the way hipFree does
    bool async_alloc = false;
    AuxStreams aux;
};

namespace {

int ctx_malloc(const tdda_ctx_s* c, void** p, size_t bytes)
{
    if (c->async_alloc) {
        if (hipMallocAsync(p, bytes, c->stream) != hipSuccess) {
            // the pool keeps freed blocks; when they are too fragmented for
            // this request, return the unused ones to the device and retry
            (void)hipGetLastError();
            hipMemPool_t pool;
            if (hipStreamSynchronize(c->stream) != hipSuccess ||
                hipDeviceGetDefaultMemPool(&pool, c->device) != hipSuccess ||
                hipMemPoolTrimTo(pool, 0) != hipSuccess || hipMallocAsync(p, bytes, c->stream) != hipSuccess) {
                (void)hipGetLastError();
                return TDDA_ERR_OOM;
            }
        }
        // usable from any stream (and by blocking copies) once the stream has reached it
        if (hipStreamSynchronize(c->stream) != hipSuccess) return TDDA_ERR_HIP;
        return TDDA_OK;
    }
    return hipMalloc(p, bytes) == hipSuccess ? TDDA_OK : TDDA_ERR_OOM;
}

void ctx_free(const tdda_ctx_s* c, void* p)
{
    if (!p) return;
    if (c->async_alloc) (void)hipFreeAsync(p, c->stream);
    else (void)hipFree(p);
}

}  // namespace

struct TilePlan {
    int pz, pz1, tfy, tfx, ps_z, ps_y, ps_x;
    int xreg;  // P3 by the register-resident kernel k_pxr
    int xwide; // k_pxr with 512 threads per workgroup (tiles up to twice as wide)
    int yreg;  // P2 / P4 by the register-resident kernel k_pyr
    int ywide; // k_pyr with 512 threads per workgroup (factors above 16)
    int fzy;   // P1 + P2 and P4 + P5 by the fused kernels k_pzy / k_pyz (tdda_zy.hpp)
    size_t lds_z, lds_z1, lds_y, lds_x, lds_zy;
};

// Upper limits on the planner's tile widths (0: none), set by the autotuner.
struct TileCaps {
    int pz = 0, pz1 = 0, tfy = 0, tfx = 0;  // pz: P5 (and P1 unless pz1), pz1: P1
    int xlds = 0;  // 1: P3 by the LDS-staged kernel even where k_pxr exists
    int xnt = 0;   // threads per workgroup of k_pxr (256 or 512; 0: the planner's choice)
    int ylds = 0;  // 1: P2 / P4 by the LDS-staged kernel even where k_pyr exists
    int nozy = 0;  // 1: separate z and y passes even where the fused kernels fit
};

struct tdda_op_s {
    tdda_ctx ctx;
    int refs = 1;
    int N, Nx, Ny, Nz, Mx, My, Mz, Hx, Hy, Hz;
    double d, lam, nback, K;
    int method = TDDA_METHOD_LDR;  // interaction tensor (A_dic_generator, AProductCore.cpp:560-571)
    int prec;
    int* cellmap = nullptr;
    int* site_xyz = nullptr;  // 3N site coordinates for the objectives, built on first use
    Cx<double>* green64 = nullptr;
    Cx<float>* green32 = nullptr;
    // full tables (Green set-up, compile-time pass kernels)
    Cx<double>*tw64[3] = {nullptr, nullptr, nullptr};
    Cx<float>*tw32[3] = {nullptr, nullptr, nullptr};
    // two-level tables (runtime-plan pass kernels)
    Cx<double>*tws64[3] = {nullptr, nullptr, nullptr};
    Cx<float>*tws32[3] = {nullptr, nullptr, nullptr};
    FftPlan pl[3];
    TilePlan tp64, tp32;
    tdda_model scratch = nullptr;
};

template <typename T>
struct VecSet {
    Cx<T>*al = nullptr, *e = nullptr, *x = nullptr, *r = nullptr, *r0 = nullptr, *p = nullptr, *ap = nullptr,
    *t = nullptr, *at = nullptr, *w1 = nullptr, *w2 = nullptr;
};

template <typename T>
struct PassGroup {
    int L = 0;                   // FFT length of the group's kernel (0: runtime plan)
    const JobDesc<T>* d = nullptr;  // descriptors of the group's jobs
    const int* ts = nullptr;     // tile starts of the group, nj + 1
    int nj = 0;
    int tiles = 0;               // tiles of the launch (upper bound of its persistent grid)
    int ncu = 0;
    size_t lds = 0;
    std::vector<int> jobs;       // the group's jobs (indices into the table)
    const int4* tmap = nullptr;  // register x / y pass: (job, ky or x, f0, nf) per tile, or none (the kernel decodes)
    const int2* zmap = nullptr;  // fused z + y passes: (job, x slab) per tile, or none
};

// Jobs of one set of launches.  Each pass runs one launch per group of jobs
// that share the pass's kernel (same compiled FFT length, or the runtime-plan
// kernel); a job's tiles, partial sums and kernel are the same in every
// table it belongs to, so its results do not depend on the other jobs.
template <typename T>
struct JobTable {
    std::vector<JobDesc<T>> h;
    JobDesc<T>* d = nullptr;
    int* tv = nullptr;  // vector-tile starts, njobs+1
    int nv = 0;
    std::vector<PassGroup<T>> gz, gz1, gy, gx;  // P5, P1, P2/P4, P3 (jobs with separate z and y passes)
    std::vector<PassGroup<T>> gzy;              // fused P1 + P2 / P4 + P5 (jobs with fzy)
    SolverState* st = nullptr;
    double* part = nullptr;
    const AuxStreams* aux = nullptr;  // the context's, for passes of several groups
    // while the solver iterates: the jobs not yet seen stopped; a pass
    // skips the groups with none (empty: every group runs)
    std::vector<char> live;
};

struct tdda_model_s {
    tdda_op op;
    int refs = 1;  // batches hold references
    int nrhs;
    size_t n3;
    VecSet<double> v64;
    VecSet<float> v32;
    JobTable<double> j64;  // FP64 solver and operator applications
    JobTable<double> jo;   // FP64 outer residual of the mixed solver
    JobTable<float> j32;   // complex64 inner solver
    SolverState* h_st = nullptr;    // pinned host mirror, 3 * nrhs
    SolverState* h_poll = nullptr;  // pinned status slots for asynchronous polling, 2 * nrhs
    hipEvent_t ev[2] = {nullptr, nullptr};
    // staging for sparse entry access
    int cap_entries = 0;
    int* d_idx = nullptr;
    Cx<double>* d_vals = nullptr;
    int* h_idx = nullptr;
    Cx<double>* h_vals = nullptr;
    std::vector<void*> allocs;
};

struct tdda_batch_s {
    tdda_ctx ctx;
    int nj = 0;
    bool has32 = false;
    std::vector<tdda_model> models;
    JobTable<double> j64, jo;
    JobTable<float> j32;
    SolverState* h_st = nullptr;
    SolverState* h_poll = nullptr;
    hipEvent_t ev[2] = {nullptr, nullptr};
    std::vector<void*> allocs;
};

namespace {

// LDS buffers of a z-pass / y-pass / x-pass tile of length L: one for the
// in-place compile-time kernels, two otherwise.
template <typename T>
int pz_buffers(int L)
{
    return pass_length_compiled<T>(0, L) && pz_inplace<T>(L) ? 1 : 2;
}
template <typename T>
int py_buffers(int L)
{
    return pass_length_compiled<T>(1, L) && py_inplace<T>(L) ? 1 : 2;
}
template <typename T>
int px_buffers(int L)
{
    return pass_length_compiled<T>(2, L) && px_inplace<T>(L) ? 1 : 2;
}

// LDS slots of an x-pass tile: twiddle table and data buffers
template <typename T>
size_t px_lds_slots(int xreg, int Mx, int tb, int tfx, int ps_x)
{
    if (xreg) return (size_t)pass_tw_slots(true, Mx, tb) + (size_t)tfx * ps_x;
    return (size_t)pass_tw_slots(pass_length_compiled<T>(2, Mx) && pass_full_twiddles<T>(2, Mx), Mx, tb) +
           (size_t)px_buffers<T>(Mx) * tfx * ps_x;
}

// LDS slots of a y-pass tile: twiddle table and data buffers
template <typename T>
size_t py_lds_slots(int yreg, int My, int tb, int tfy, int ps_y)
{
    if (yreg) return (size_t)pass_tw_slots(true, My, tb) + (size_t)tfy * ps_y;
    return (size_t)pass_tw_slots(pass_length_compiled<T>(1, My) && pass_full_twiddles<T>(1, My), My, tb) +
           (size_t)py_buffers<T>(My) * tfy * ps_y;
}

// Tile widths of the passes for element type T: a tile holds at most
// `cap` elements (pencils x length, two LDS buffers of it), rows of the
// strided passes are kept to whole 128-byte lines where the budget allows,
// and the Green product of the x pass covers at most 8 points per thread.
template <typename T>
int plan_tiles(const tdda_op_s* op, TilePlan& t, const TileCaps& caps = TileCaps())
{
    const size_t s = sizeof(Cx<T>);
    // prefetch capacity of a workgroup in each pass (elements), as compiled
    // into the pass kernel that will run: ept_len of its FFT length
    auto full = [&](int kind, int L) { return kThreads * ept_len<T>(pass_length_compiled<T>(kind, L) ? L : 0); };
    const bool ipx = px_buffers<T>(op->Mx) == 1;  // in-place x pass: loaded rows only
    const bool ipy = py_buffers<T>(op->My) == 1;  // in-place y pass: rows read only
    const int fz = full(0, op->Mz);
    const int fy = ipy ? kThreads * py_ept<T, true>(op->My) : full(1, op->My);
    const int fx = ipx ? kThreads * px_ept<T>(op->Mx) : full(2, op->Mx);
    const int env = env_int(sizeof(T) == 8 ? "TDDA_TILE_C128" : "TDDA_TILE_C64", 1 << 30);
    const int capz = std::min(fz, env), capy = std::min(fy, env), capx = std::min(fx, env);
    const int cap1 = fz / 3;  // P1: up to three input vectors
    const int unit = (int)(128 / s);
    t.ps_z = lds_pencil_stride<T>(op->Mz);
    t.ps_y = lds_pencil_stride<T>(op->My);
    t.ps_x = lds_pencil_stride<T>(op->Mx);
    // one tile must hold: 3 z pencils (P5 reads all kz), one y column of My
    // rows (P4), and the Nx loaded rows of 3 x columns (P3; the Mx - Nx
    // padding rows are zero-filled in LDS, bounded by the LDS check below)
    if (3 * op->Mz > capz || op->My > capy || 3 * op->Nx > capx)
        return fail(TDDA_ERR_UNSUPPORTED, "FFT length too long for the pass tiles of this precision");
    t.pz = std::max(1, std::min({capz / (3 * op->Mz), 64, op->Nx * op->Ny}));
    if (3 * op->Nz > cap1) return fail(TDDA_ERR_UNSUPPORTED, "Nz too large for the z-pass tiles of this precision");
    t.pz1 = std::max(1, std::min({cap1 / (3 * op->Nz), capz / (3 * op->Mz), 64, op->Nx * op->Ny}));
    int ty = std::min(capy / op->My, 3 * op->Mz);
    if (ipy) ty = std::min(ty, py_inplace_pencils<T>());
    else if (ty >= unit && ty < 3 * op->Mz) ty = (ty / unit) * unit;
    t.tfy = std::max(1, ty);
    int tx;
    if (ipx) {
        tx = std::min({px_np<T>(op->Mx), 3 * op->Mz, capx / op->Nx, 3 * (8 * kThreads / op->Mx)});
        tx = (tx / 3) * 3;
    } else {
        tx = std::min({capx / op->Mx, 3 * op->Mz, 3 * (8 * kThreads / op->Mx)});
        tx = (tx / 3) * 3;
        if (tx >= 3 * unit && tx < 3 * op->Mz) tx = (tx / (3 * unit)) * (3 * unit);
    }
    t.tfx = std::max(3, tx);
    if (caps.pz > 0) {
        t.pz = std::min(t.pz, caps.pz);
        t.pz1 = std::min(t.pz1, caps.pz);
    }
    if (caps.pz1 > 0) t.pz1 = std::min(t.pz1, caps.pz1);
    if (caps.tfy > 0) t.tfy = std::min(t.tfy, caps.tfy);
    if (caps.tfx > 0) t.tfx = std::max(3, std::min(t.tfx, (caps.tfx / 3) * 3));
    // register-resident x pass where its kernel exists (TDDA_XPASS=0 or the
    // autotuner's xlds cap keep the LDS-staged kernel): tiles of at most
    // pxr_nf columns, odd pencil stride, full twiddle table
    t.xreg = pass_length_compiled<T>(3, op->Mx) && !caps.xlds && env_int("TDDA_XPASS", 1) != 0;
    t.xwide = 0;
    if (t.xreg) {
        // 256 threads unless the autotuner measured the 512-thread kernel
        // (tiles up to twice as wide) faster for this operator, or
        // TDDA_XNT=512 is set; untuned operators keep one kernel per length,
        // so a batch of them needs one x-pass launch per length
        const int env_nt = env_int("TDDA_XNT", 0);
        // (lengths with a factor above 16 always take 512: 256 threads
        // would give tiles of 6 columns)
        const int nt = caps.xnt ? caps.xnt : (env_nt == 512 || pxr_l1(op->Mx) > 16) ? 512 : kThreads;
        t.xwide = nt == 512;
        t.tfx = std::max(3, std::min(pxr_nf<T>(op->Mx, nt), 3 * op->Mz));
        if (caps.tfx > 0) t.tfx = std::max(3, std::min(t.tfx, (caps.tfx / 3) * 3));
        t.ps_x = op->Mx | 1;
    }
    if (!t.xreg && (t.tfx / 3) * op->Mx > 8 * kThreads) return fail(TDDA_ERR_UNSUPPORTED, "x pass tile too wide");
    // register-resident y passes likewise (TDDA_YPASS=0 or the ylds cap keep
    // the LDS-staged kernels): tiles of at most pyr_nf columns
    t.yreg = pass_length_compiled<T>(3, op->My) && !caps.ylds && env_int("TDDA_YPASS", 1) != 0;
    t.ywide = 0;
    if (t.yreg) {
        t.ywide = pxr_l1(op->My) > 16;  // 512 threads: rows of 16 columns instead of 8
        t.tfy = std::max(1, std::min(pyr_nf<T>(op->My, t.ywide ? 512 : kThreads), 3 * op->Mz));
        if (caps.tfy > 0) t.tfy = std::min(t.tfy, caps.tfy);
        t.ps_y = op->My | 1;
    }
    auto tws = [&](int kind, int L, int tb) { return pass_tw_slots(pass_length_compiled<T>(kind, L) && pass_full_twiddles<T>(kind, L), L, tb); };
    t.lds_z = (size_t)(tws(0, op->Mz, op->pl[2].tb) + pz_buffers<T>(op->Mz) * 3 * t.pz * t.ps_z) * s;
    t.lds_z1 = (size_t)(tws(0, op->Mz, op->pl[2].tb) + pz_buffers<T>(op->Mz) * 3 * t.pz1 * t.ps_z) * s;
    t.lds_y = py_lds_slots<T>(t.yreg, op->My, op->pl[1].tb, t.tfy, t.ps_y) * s;
    t.lds_x = px_lds_slots<T>(t.xreg, op->Mx, op->pl[0].tb, t.tfx, t.ps_x) * s;
    if (t.xwide && t.lds_x > op->ctx->lds_budget) {  // a wide tile too large: 256 threads
        t.xwide = 0;
        t.tfx = std::max(3, std::min(pxr_nf<T>(op->Mx), 3 * op->Mz));
        if (caps.tfx > 0) t.tfx = std::max(3, std::min(t.tfx, (caps.tfx / 3) * 3));
        t.lds_x = px_lds_slots<T>(t.xreg, op->Mx, op->pl[0].tb, t.tfx, t.ps_x) * s;
    }
    // fused z + y passes where their kernels exist for (My, Mz) and a
    // whole x slab fits in LDS with at least two workgroups per CU
    // (TDDA_FUSEZY=0 or the autotuner's nozy cap keep the separate passes)
    t.fzy = 0;
    t.lds_zy = 0;
    if (t.yreg && !t.ywide && zy_ylen_ok(op->My) && zy_zlen_ok<T>(op->Mz) && !caps.nozy &&
        env_int("TDDA_FUSEZY", 1) != 0) {
        const size_t lz = zy_lds_slots(op->My, op->Ny, op->Nz, op->Mz, t.tfy, t.ps_y) * s;
        if (lds_resident_workgroups(lz + 512) >= 2 && lz <= op->ctx->lds_budget) {
            t.fzy = 1;
            t.lds_zy = lz;
            // one z tile per slab: the fused inverse keeps one partial-sum
            // slot per slab (k_finish counts Nx Ny / pz of them)
            t.pz = t.pz1 = op->Ny;
            t.lds_z = t.lds_z1 = 0;
        }
    }
    if (std::max({t.lds_z, t.lds_z1, t.lds_y, t.lds_x, t.lds_zy}) > op->ctx->lds_budget)
        return fail(TDDA_ERR_UNSUPPORTED, "pass tile exceeds the LDS budget");
    return TDDA_OK;
}

template <typename T>
int dev_alloc(const tdda_ctx_s* c, std::vector<void*>& arena, T** p, size_t count)
{
    void* q = nullptr;
    if (count == 0) count = 1;
    if (ctx_malloc(c, &q, count * sizeof(T)) != TDDA_OK)
        return fail(TDDA_ERR_OOM, "device allocation of " + std::to_string(count * sizeof(T)) + " bytes failed");
    arena.push_back(q);
    *p = reinterpret_cast<T*>(q);
    return TDDA_OK;
}

template <typename T>
int alloc_vecs(tdda_model_s* m, VecSet<T>& v)
{
    const size_t n = m->n3 * m->nrhs;
    const tdda_op_s* op = m->op;
    std::vector<void*>& a = m->allocs;
    CHK(dev_alloc(op->ctx, a, &v.al, n));
    CHK(dev_alloc(op->ctx, a, &v.e, n));
    CHK(dev_alloc(op->ctx, a, &v.x, n));
    CHK(dev_alloc(op->ctx, a, &v.r, n));
    CHK(dev_alloc(op->ctx, a, &v.r0, n));
    CHK(dev_alloc(op->ctx, a, &v.p, n));
    CHK(dev_alloc(op->ctx, a, &v.ap, n));
    CHK(dev_alloc(op->ctx, a, &v.t, n));
    CHK(dev_alloc(op->ctx, a, &v.at, n));
    CHK(dev_alloc(op->ctx, a, &v.w1, (size_t)op->Nx * op->Ny * op->Mz * 3 * m->nrhs));
    CHK(dev_alloc(op->ctx, a, &v.w2, (size_t)op->Nx * op->My * op->Mz * 3 * m->nrhs));
    HIPCHK(hipMemsetAsync(v.x, 0, n * sizeof(Cx<T>), op->ctx->stream));
    HIPCHK(hipMemsetAsync(v.al, 0, n * sizeof(Cx<T>), op->ctx->stream));
    HIPCHK(hipMemsetAsync(v.e, 0, n * sizeof(Cx<T>), op->ctx->stream));
    return TDDA_OK;
}

// Host job descriptors of one model for precision T (one job per right-hand
// side).  r_vec / r0_vec let the outer residual table of the mixed solver
// write its residual and r0 copy into scratch vectors.
template <typename T>
void fill_jobs(const tdda_op_s* op, int nrhs, std::vector<JobDesc<T>>& out, const VecSet<T>& v, const Cx<T>* green,
               Cx<T>* const* tw, const TilePlan& tp, Cx<T>* r_vec, Cx<T>* r0_vec)
{
    const size_t n3 = 3 * (size_t)op->N;
    const size_t w1n = (size_t)op->Nx * op->Ny * op->Mz * 3;
    const size_t w2n = (size_t)op->Nx * op->My * op->Mz * 3;
    for (int j = 0; j < nrhs; j++) {
        JobDesc<T> J;
        memset(&J, 0, sizeof(J));
        J.Nx = op->Nx;
        J.Ny = op->Ny;
        J.Nz = op->Nz;
        J.Mx = op->Mx;
        J.My = op->My;
        J.Mz = op->Mz;
        J.Hx = op->Hx;
        J.Hy = op->Hy;
        J.Hz = op->Hz;
        J.N = op->N;
        J.pz = tp.pz;
        J.pz1 = tp.pz1;
        J.tfy = tp.tfy;
        J.tfx = tp.tfx;
        J.nfy = (3 * op->Mz + tp.tfy - 1) / tp.tfy;
        J.nfx = (3 * op->Mz + tp.tfx - 1) / tp.tfx;
        J.ps_z = tp.ps_z;
        J.ps_y = tp.ps_y;
        J.ps_x = tp.ps_x;
        J.xreg = tp.xreg;
        J.xwide = tp.xwide;
        J.yreg = tp.yreg;
        J.ywide = tp.ywide;
        J.fzy = tp.fzy;
        J.gidx = j;
        J.gsize = nrhs;
        J.cellmap = op->cellmap;
        J.green = green;
        J.twx = tw[0];
        J.twy = tw[1];
        J.twz = tw[2];
        J.twfx = tw[3];
        J.twfy = tw[4];
        J.twfz = tw[5];
        J.plx = op->pl[0];
        J.ply = op->pl[1];
        J.plz = op->pl[2];
        J.w1 = v.w1 + j * w1n;
        J.w2 = v.w2 + j * w2n;
        J.in = v.x + j * n3;
        J.out = v.ap + j * n3;
        J.al = v.al + j * n3;
        J.x = v.x + j * n3;
        J.r = r_vec + j * n3;
        J.r0 = r0_vec + j * n3;
        J.p = v.p + j * n3;
        J.ap = v.ap + j * n3;
        J.t = v.t + j * n3;
        J.at = v.at + j * n3;
        J.e = v.e + j * n3;
        out.push_back(J);
    }
}

struct JobTiles {
    int z, z1, y, x, v;
};

template <typename T>
JobTiles job_tiles(const JobDesc<T>& J)
{
    JobTiles t;
    t.z = (J.Nx * J.Ny + J.pz - 1) / J.pz;
    t.z1 = (J.Nx * J.Ny + J.pz1 - 1) / J.pz1;
    t.y = J.Nx * J.nfy;
    t.x = J.My * J.nfx;
    t.v = (int)((3 * (size_t)J.N + kThreads * kVecPer - 1) / (kThreads * kVecPer));
    return t;
}

// Per-pass launch groups of a job table whose host descriptors are final:
// device descriptor array, vector-tile prefix sums and, for every pass kind,
// one group per kernel (compiled FFT length, or 0 for the runtime-plan
// kernel) with its own descriptor subset, tile prefix sums, LDS size and
// persistent grid.  `put(host, count)` copies an array to device memory and
// returns its device address (nullptr on failure).
// Tile map of an x-pass launch group: per tile the job (within the
// group), ky, first column and column count, by the arithmetic of s_tile
// (interleaved right-hand sides of one operator, mirrored ky order).
template <typename T>
std::vector<int4> ypass_tile_map(const std::vector<JobDesc<T>>& sub, const std::vector<int>& ts)
{
    std::vector<int4> m(ts.back());
    int j = 0;
    for (int r = 0; r < ts.back(); r++) {
        while (ts[j + 1] <= r) j++;
        const JobDesc<T>& J = sub[j];
        int x, fb;
        ytile_map(r - ts[j], J.Nx, J.nfy, x, fb);
        const int f0 = fb * J.tfy;
        m[r] = make_int4(j, x, f0, std::min(J.tfy, J.Mz * 3 - f0));
    }
    return m;
}

template <typename T>
std::vector<int4> xpass_tile_map(const std::vector<JobDesc<T>>& sub, const std::vector<int>& ts)
{
    std::vector<int4> m(ts.back());
    int j0 = 0;
    for (int r = 0; r < ts.back(); r++) {
        while (ts[j0 + 1] <= r) j0++;
        const JobDesc<T>& J0 = sub[j0];
        const int first = j0 - J0.gidx;
        int ky, fb, k;
        xtile_map(r - ts[first], J0.My, J0.nfx, J0.gsize, ky, fb, k);
        const int job = first + k;
        const int f0 = fb * J0.tfx;
        m[r] = make_int4(job, ky, f0, std::min(J0.tfx, sub[job].Mz * 3 - f0));
    }
    return m;
}

// Bytes of the tile maps of a table (build_groups with xmap): register x
// and y passes and fused z + y passes.
template <typename T>
size_t xmap_bound(const JobTable<T>& jt)
{
    size_t n = 0;
    size_t z = 0;
    for (const JobDesc<T>& J : jt.h) {
        n += (size_t)J.My * J.nfx + (size_t)J.Nx * J.nfy;
        z += (size_t)J.Nx;
    }
    return n * sizeof(int4) + z * sizeof(int2) + 512 * (jt.h.size() + 1);
}

template <typename T, typename Put>
int build_groups(const tdda_ctx_s* c, JobTable<T>& jt, Put&& put, bool xmap = false)
{
    const int nj = (int)jt.h.size();
    const size_t s = sizeof(Cx<T>);
    std::vector<JobTiles> tiles(nj);
    std::vector<int> hv(nj + 1, 0);
    for (int j = 0; j < nj; j++) {
        tiles[j] = job_tiles(jt.h[j]);
        hv[j + 1] = hv[j] + tiles[j].v;
    }
    jt.d = const_cast<JobDesc<T>*>(put(jt.h.data(), (size_t)nj));
    jt.tv = const_cast<int*>(put(hv.data(), (size_t)nj + 1));
    if (!jt.d || !jt.tv) return fail(TDDA_ERR_OOM, "job table upload");
    jt.nv = hv[nj];
    jt.aux = &c->aux;
    // per pass kind: FFT length, pass kind of pass_length_compiled, tiles, LDS bytes
    // kind: 0 z, 1 y, 2 x, 4 fused z + y; `fused`: the jobs the groups take
    // (0: separate z and y passes, 1: fused, 2: all)
    auto build = [&](std::vector<PassGroup<T>>& groups, int kind, int fused, auto len, auto ntiles,
                     auto lds) -> int {
        groups.clear();
        std::vector<int> key(nj);
        for (int j = 0; j < nj; j++) {
            const int L = len(jt.h[j]);
            const bool reg = (kind == 2 && jt.h[j].xreg) || (kind == 1 && jt.h[j].yreg);
            const bool wide = (kind == 2 && jt.h[j].xwide) || (kind == 1 && jt.h[j].ywide);
            key[j] = kind == 4 ? L
                     : reg     ? (L | kPxrKey | (wide ? kPxrWide : 0))
                     : pass_length_compiled<T>(kind, L) ? L : 0;
            if (fused != 2 && (jt.h[j].fzy != 0) != (fused == 1)) key[j] = -1;  // not in these groups
        }
        std::vector<int> keys;
        for (int j = 0; j < nj; j++)
            if (key[j] >= 0 && std::find(keys.begin(), keys.end(), key[j]) == keys.end()) keys.push_back(key[j]);
        for (int k : keys) {
            PassGroup<T> g;
            g.L = k;
            std::vector<JobDesc<T>> sub;
            std::vector<int> subt;
            for (int j = 0; j < nj; j++) {
                if (key[j] != k) continue;
                sub.push_back(jt.h[j]);
                g.jobs.push_back(j);
                subt.push_back(ntiles(tiles[j]));
                g.lds = std::max(g.lds, lds(jt.h[j]));
            }
            if (g.lds > c->lds_budget) return fail(TDDA_ERR_UNSUPPORTED, "pass tile exceeds the LDS budget");
            std::vector<int> ts(1, 0);
            for (int t : subt) ts.push_back(ts.back() + t);
            g.nj = (int)sub.size();
            g.tiles = ts.back();
            g.ncu = c->ncu;
            g.d = (sub.size() == (size_t)nj) ? jt.d : put(sub.data(), sub.size());
            g.ts = put(ts.data(), ts.size());
            if (!g.d || !g.ts) return fail(TDDA_ERR_OOM, "job table upload");
            if (xmap && (kind == 1 || kind == 2) && (k & kPxrKey)) {
                const std::vector<int4> tm = kind == 2 ? xpass_tile_map(sub, ts) : ypass_tile_map(sub, ts);
                g.tmap = put(tm.data(), tm.size());
                if (!g.tmap) return fail(TDDA_ERR_OOM, "job table upload");
            }
            if (xmap && kind == 4) {
                std::vector<int2> zm(ts.back());
                for (size_t q = 0; q + 1 < ts.size(); q++)
                    for (int r = ts[q]; r < ts[q + 1]; r++) zm[r] = make_int2((int)q, r - ts[q]);
                g.zmap = put(zm.data(), zm.size());
                if (!g.zmap) return fail(TDDA_ERR_OOM, "job table upload");
            }
            groups.push_back(g);
        }
        return TDDA_OK;
    };
    auto Lz = [](const JobDesc<T>& J) { return J.Mz; };
    auto tws = [](int kind, int L, int tb) { return pass_tw_slots(pass_length_compiled<T>(kind, L) && pass_full_twiddles<T>(kind, L), L, tb); };
    CHK(build(jt.gz, 0, 0, Lz, [](const JobTiles& t) { return t.z; },
              [&](const JobDesc<T>& J) { return (size_t)(tws(0, J.Mz, J.plz.tb) + pz_buffers<T>(J.Mz) * 3 * J.pz * J.ps_z) * s; }));
    CHK(build(jt.gz1, 0, 0, Lz, [](const JobTiles& t) { return t.z1; },
              [&](const JobDesc<T>& J) { return (size_t)(tws(0, J.Mz, J.plz.tb) + pz_buffers<T>(J.Mz) * 3 * J.pz1 * J.ps_z) * s; }));
    CHK(build(jt.gy, 1, 0, [](const JobDesc<T>& J) { return J.My; }, [](const JobTiles& t) { return t.y; },
              [&](const JobDesc<T>& J) { return py_lds_slots<T>(J.yreg, J.My, J.ply.tb, J.tfy, J.ps_y) * s; }));
    CHK(build(jt.gx, 2, 2, [](const JobDesc<T>& J) { return J.Mx; }, [](const JobTiles& t) { return t.x; },
              [&](const JobDesc<T>& J) { return px_lds_slots<T>(J.xreg, J.Mx, J.plx.tb, J.tfx, J.ps_x) * s; }));
    // fused z + y: one tile per x slab
    CHK(build(jt.gzy, 4, 1, [](const JobDesc<T>& J) { return J.My; }, [](const JobTiles& t) { return t.z; },
              [&](const JobDesc<T>& J) { return zy_lds_slots(J.My, J.Ny, J.Nz, J.Mz, J.tfy, J.ps_y) * s; }));
    return TDDA_OK;
}

// Partial-sum slots of one job (deterministic reductions: one slot of 4
// values per tile of its largest reducing pass).
template <typename T>
size_t job_part_size(const JobDesc<T>& J)
{
    const JobTiles t = job_tiles(J);
    return kPart * (size_t)std::max(t.z, t.v);
}

// Upper bound of the bytes build_groups uploads for a table of nj jobs
// (descriptors: the table's own and the group subsets, each job in at most
// four groups; tile starts; 256-byte alignment of every array).
template <typename T>
size_t table_bound(size_t nj)
{
    return 5 * nj * sizeof(JobDesc<T>) + (11 * nj + 16) * sizeof(int) + 256 * (2 + 9 * nj) + 1024;
}

// Completes a job table from its host descriptors: per-job solver state,
// partial-sum buffers and counters, then the launch groups, all in device
// memory owned by `arena`.
template <typename T>
int finalize_table(std::vector<void*>& arena, const tdda_ctx_s* c, JobTable<T>& jt)
{
    const int nj = (int)jt.h.size();
    std::vector<size_t> poff(nj + 1, 0);
    for (int j = 0; j < nj; j++) poff[j + 1] = poff[j] + job_part_size(jt.h[j]);
    CHK(dev_alloc(c, arena, &jt.st, nj));
    CHK(dev_alloc(c, arena, &jt.part, poff[nj]));
    HIPCHK(hipMemsetAsync(jt.st, 0, sizeof(SolverState) * nj, c->stream));
    for (int j = 0; j < nj; j++) {
        jt.h[j].st = jt.st + j;
        jt.h[j].part = jt.part + poff[j];
    }
    // the descriptor and tile arrays of all launch groups in one device
    // block, filled from one host image with one copy (a request of the
    // serving loop builds three tables: one allocation each, not dozens)
    const size_t cap = table_bound<T>((size_t)nj) + xmap_bound(jt);
    char* dblk = nullptr;
    CHK(dev_alloc(c, arena, &dblk, cap));
    std::vector<char> img(cap);
    size_t used = 0;
    auto put = [&](const auto* host, size_t n) {
        using U = std::remove_const_t<std::remove_pointer_t<decltype(host)>>;
        const size_t off = (used + 255) & ~(size_t)255, bytes = sizeof(U) * n;
        if (off + bytes > cap) return (const U*)nullptr;
        memcpy(img.data() + off, host, bytes);
        used = off + bytes;
        return reinterpret_cast<const U*>(dblk + off);
    };
    CHK(build_groups(c, jt, put, true));
    HIPCHK(hipMemcpy(dblk, img.data(), used, hipMemcpyHostToDevice));
    return TDDA_OK;
}

// Launches the groups of one pass: one group on the stream itself; several
// spread over the stream and the context's auxiliary streams (largest
// groups first, round robin), forked from and joined back into the stream,
// so the pass is complete on the stream when the next work there starts.
//
// A pass may take groups of two lists (gb may be empty): the separate y or
// z pass of some jobs and the fused z + y pass of the others, which run
// side by side.  Groups whose jobs have all stopped (jt.live) are not
// launched: their kernels would return at once on every tile.
template <typename T, typename LA, typename LB>
void fan_out2(const JobTable<T>& jt, const std::vector<PassGroup<T>>& ga, LA&& launch_a,
              const std::vector<PassGroup<T>>& gb, LB&& launch_b, hipStream_t s)
{
    const AuxStreams* a = jt.aux;
    auto group = [&](size_t i) -> const PassGroup<T>& { return i < ga.size() ? ga[i] : gb[i - ga.size()]; };
    auto run = [&](size_t i, hipStream_t st) {
        if (i < ga.size()) launch_a(ga[i], st);
        else launch_b(gb[i - ga.size()], st);
    };
    std::vector<size_t> order;
    order.reserve(ga.size() + gb.size());
    for (size_t i = 0; i < ga.size() + gb.size(); i++) {
        bool any = jt.live.empty();
        for (size_t k = 0; !any && k < group(i).jobs.size(); k++) any = jt.live[group(i).jobs[k]] != 0;
        if (any) order.push_back(i);
    }
    const size_t n = order.size();
    if (n <= 1 || !a || a->n == 0) {
        for (size_t i : order) run(i, s);
        return;
    }
    std::stable_sort(order.begin(), order.end(), [&](size_t x, size_t y) { return group(x).tiles > group(y).tiles; });
    const int nstreams = std::min<int>(a->n + 1, (int)n);
    (void)hipEventRecord(a->fork, s);
    for (int k = 1; k < nstreams; k++) (void)hipStreamWaitEvent(a->s[k - 1], a->fork, 0);
    for (size_t i = 0; i < order.size(); i++) {
        const int k = (int)(i % nstreams);
        run(order[i], k == 0 ? s : a->s[k - 1]);
    }
    for (int k = 1; k < nstreams; k++) {
        (void)hipEventRecord(a->join[k - 1], a->s[k - 1]);
        (void)hipStreamWaitEvent(s, a->join[k - 1], 0);
    }
}

template <typename T, typename L>
void fan_out(const JobTable<T>& jt, const std::vector<PassGroup<T>>& gs, hipStream_t s, L&& launch)
{
    static const std::vector<PassGroup<T>> none;
    fan_out2(jt, gs, launch, none, launch, s);
}

template <typename T>
void run_p1(const JobTable<T>& jt, int mode, hipStream_t s)
{
    fan_out(jt, jt.gz1, s, [&](const PassGroup<T>& g, hipStream_t st) {
        launch_p1<T>(g.L, mode, PassLaunch{dim3(g.tiles), g.lds, st, g.ncu}, g.d, g.ts, g.nj);
    });
}

// P5 of the jobs with separate passes, and P4 + P5 of the fused jobs
template <typename T>
void run_p5(const JobTable<T>& jt, int mode, int diag, hipStream_t s)
{
    fan_out2(
        jt, jt.gz,
        [&](const PassGroup<T>& g, hipStream_t st) {
            launch_p5<T>(g.L, mode, diag, PassLaunch{dim3(g.tiles), g.lds, st, g.ncu}, g.d, g.ts, g.nj);
        },
        jt.gzy,
        [&](const PassGroup<T>& g, hipStream_t st) {
            launch_pyz<T>(g.L, mode, mode == OUT_PLAIN ? diag : 1, PassLaunch{dim3(g.tiles), g.lds, st, g.ncu}, g.d,
                          g.ts, g.nj, g.zmap);
        },
        s);
}

// P2 / P4 of the jobs with separate passes; forward (P2) also runs P1 + P2
// of the fused jobs (`inmode`: their input mode, as P1's)
template <typename T>
void run_py(const JobTable<T>& jt, bool inv, int solver, hipStream_t s, int inmode = IN_PLAIN)
{
    static const std::vector<PassGroup<T>> none;
    fan_out2(
        jt, jt.gy,
        [&](const PassGroup<T>& g, hipStream_t st) {
            launch_py<T>(g.L, inv, solver, PassLaunch{dim3(g.tiles), g.lds, st, g.ncu}, g.d, g.ts, g.nj, g.tmap);
        },
        inv ? none : jt.gzy,
        [&](const PassGroup<T>& g, hipStream_t st) {
            launch_pzy<T>(g.L, inmode, PassLaunch{dim3(g.tiles), g.lds, st, g.ncu}, g.d, g.ts, g.nj, g.zmap);
        },
        s);
}

template <typename T>
void run_px(const JobTable<T>& jt, int solver, hipStream_t s)
{
    fan_out(jt, jt.gx, s, [&](const PassGroup<T>& g, hipStream_t st) {
        launch_px<T>(g.L, solver, PassLaunch{dim3(g.tiles), g.lds, st, g.ncu}, g.d, g.ts, g.nj, g.tmap);
    });
}

// one operator application: P1 .. P5
template <typename T>
void launch_apply(const JobTable<T>& jt, int nj, int inmode, int outmode, int diag, int solver, hipStream_t s)
{
    run_p1<T>(jt, inmode, s);
    run_py<T>(jt, false, solver, s, inmode);
    run_px<T>(jt, solver, s);
    run_py<T>(jt, true, solver, s);
    run_p5<T>(jt, outmode, diag, s);
    switch (outmode) {
    case OUT_RES: k_finish<T, FIN_RES><<<dim3(nj), dim3(kThreads), 0, s>>>(jt.d, nj); break;
    case OUT_AP: k_finish<T, FIN_AP><<<dim3(nj), dim3(kThreads), 0, s>>>(jt.d, nj); break;
    case OUT_AT: k_finish<T, FIN_AT><<<dim3(nj), dim3(kThreads), 0, s>>>(jt.d, nj); break;
    default: break;
    }
}

// one BiCGSTAB iteration (or the unrolled first step when state.first):
// two operator applications and their finishing launches; the iteration's
// updates of x and r are applied by the next iteration's P1 (or k_settle)
template <typename T>
void launch_iteration(const JobTable<T>& jt, int nj, hipStream_t s)
{
    launch_apply<T>(jt, nj, IN_P, OUT_AP, 1, 1, s);
    launch_apply<T>(jt, nj, IN_T, OUT_AT, 1, 1, s);
}

// x and r of the jobs that have stopped with their last update pending
// (after the iterations of a solve, before x or r is read)
template <typename T>
void launch_settle(const JobTable<T>& jt, int nj, hipStream_t s)
{
    k_settle<T><<<dim3(jt.nv), dim3(kThreads), 0, s>>>(jt.d, jt.tv, nj);
    k_finish<T, FIN_SETTLE><<<dim3(nj), dim3(kThreads), 0, s>>>(jt.d, nj);
}

// right Jacobi preconditioning: P = D^-1 y of the settled jobs that have
// stopped (DIV), or y = D P before iterating from a given P (!DIV)
template <typename T, bool DIV>
void launch_prec_x(const JobTable<T>& jt, int nj, hipStream_t s)
{
    k_prec_x<T, DIV><<<dim3(jt.nv), dim3(kThreads), 0, s>>>(jt.d, jt.tv, nj);
}

// The job tables and host buffers one solve runs on: a model's own, or the
// concatenation of several models' (a batch).
struct SolveSet {
    hipStream_t s;
    int nj;
    JobTable<double>* j64;
    JobTable<double>* jo;
    JobTable<float>* j32;  // nullptr without complex64 storage
    SolverState* h_st;     // pinned, 3 * nj
    SolverState* h_poll;   // pinned, 2 * nj
    hipEvent_t* ev;        // 2 events
};

int check_launch()
{
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(TDDA_ERR_HIP, std::string("kernel launch: ") + hipGetErrorString(e));
    return TDDA_OK;
}

// Runs BiCGSTAB iterations for every running job until each has stopped.
// States must already be initialised (first = 1) and r, r0, x, rho set.
// Chunks of iterations are queued ahead of the status they depend on: while
// one chunk runs, the host reads the status copied after the previous chunk,
// so the device does not idle between chunks.  The size of the next chunk
// comes from each job's observed residual decay; chunks queued after every
// job has stopped run as no-ops (each kernel returns on the stop flag).
template <typename T>
int run_iterations(SolveSet& S, JobTable<T>& jt, SolverState* hst, int max_total, int check_every)
{
    hipStream_t s = S.s;
    const int nj = S.nj;
    SolverState* slot[2] = {S.h_poll, S.h_poll + nj};
    bool busy[2] = {false, false};
    int upto[2] = {0, 0};  // iterations (including the first step) issued when slot k was filled
    int issued = 0;
    const int limit = max_total + 1;
    auto issue = [&](int k, int chunk) -> int {
        chunk = std::max(1, std::min(chunk, limit - issued));
        for (int i = 0; i < chunk; i++) launch_iteration<T>(jt, nj, s);
        issued += chunk;
        HIPCHK(hipMemcpyAsync(slot[k], jt.st, sizeof(SolverState) * nj, hipMemcpyDeviceToHost, s));
        HIPCHK(hipEventRecord(S.ev[k], s));
        busy[k] = true;
        upto[k] = issued;
        return check_launch();
    };
    std::vector<double> prev_err(nj, -1.0);
    std::vector<int> prev_it(nj, 0);
    // jobs seen stopped leave the launch groups; the table runs every group
    // again once the iterations end
    jt.live.assign(nj, 1);
    struct LiveReset {
        std::vector<char>& v;
        ~LiveReset() { v.clear(); }
    } live_reset{jt.live};
    int cur = 0;
    CHK(issue(0, check_every > 0 ? check_every : 4));
    if (issued < limit) CHK(issue(1, check_every > 0 ? check_every : 4));
    while (true) {
        HIPCHK(hipEventSynchronize(S.ev[cur]));
        busy[cur] = false;
        const SolverState* St = slot[cur];
        bool running = false;
        int need = 0;
        for (int j = 0; j < nj; j++) {
            if (St[j].status != ST_RUNNING) {
                jt.live[j] = 0;
                continue;
            }
            running = true;
            int rem = 4;
            if (prev_err[j] > 0 && St[j].iter > prev_it[j] && St[j].err < prev_err[j] && St[j].err > 0) {
                const double rate = std::pow(St[j].err / prev_err[j], 1.0 / (St[j].iter - prev_it[j]));
                if (rate < 0.999) rem = (int)std::ceil(std::log(St[j].tol / St[j].err) / std::log(rate));
            }
            rem = std::max(1, std::min(rem, 64));
            prev_err[j] = St[j].err;
            prev_it[j] = St[j].iter;
            need = std::max(need, rem);
        }
        if (!running || issued >= limit) {
            if (!running || !busy[1 - cur]) {
                memcpy(hst, St, sizeof(SolverState) * nj);
                if (!running) return TDDA_OK;
                // iteration cap reached with jobs still running: the device
                // marks them ST_MAXIT in the last queued chunk
                HIPCHK(hipStreamSynchronize(s));
                HIPCHK(hipMemcpy(hst, jt.st, sizeof(SolverState) * nj, hipMemcpyDeviceToHost));
                return TDDA_OK;
            }
            cur = 1 - cur;
            continue;
        }
        const int in_flight = busy[1 - cur] ? upto[1 - cur] - upto[cur] : 0;
        const int more = check_every > 0 ? check_every : std::min(32, need - in_flight);
        if (more > 0 || !busy[1 - cur]) CHK(issue(cur, std::max(more, 1)));
        if (busy[1 - cur]) cur = 1 - cur;
    }
}

template <typename T>
int init_states(SolveSet& S, JobTable<T>& jt, SolverState* hst, const std::vector<int>& active, double tol,
                int maxit, const std::vector<double>* tols = nullptr, int prec = 0)
{
    for (int j = 0; j < S.nj; j++) {
        SolverState& St = hst[j];
        memset(&St, 0, sizeof(St));
        St.tol = tols ? (*tols)[j] : tol;
        St.maxit = maxit;
        St.first = 1;
        St.prec = prec;
        St.idle = !active[j];
        St.status = active[j] ? ST_RUNNING : ST_CONVERGED;
    }
    HIPCHK(hipMemcpyAsync(jt.st, hst, sizeof(SolverState) * S.nj, hipMemcpyHostToDevice, S.s));
    return TDDA_OK;
}

template <typename T>
void launch_init_zero(const JobTable<T>& jt, int nj, hipStream_t s)
{
    k_init_zero<T><<<dim3(jt.nv), dim3(kThreads), 0, s>>>(jt.d, jt.tv, nj);
    k_finish<T, FIN_INIT><<<dim3(nj), dim3(kThreads), 0, s>>>(jt.d, nj);
}

int solve_fp64(SolveSet& S, const tdda_solve_opts* o, tdda_solve_info* info)
{
    const int nj = S.nj;
    SolverState* hst = S.h_st;
    std::vector<int> active(nj, 1);
    const int prec = o->precond == TDDA_PRECOND_JACOBI;
    CHK(init_states<double>(S, *S.j64, hst, active, o->tol, o->max_it, nullptr, prec));
    if (o->x0_zero) {
        launch_init_zero<double>(*S.j64, nj, S.s);
    } else {
        launch_apply<double>(*S.j64, nj, IN_X, OUT_RES, 1, 1, S.s);
        if (prec) launch_prec_x<double, false>(*S.j64, nj, S.s);
    }
    CHK(check_launch());
    CHK(run_iterations<double>(S, *S.j64, hst, o->max_it, o->check_every));
    launch_settle<double>(*S.j64, nj, S.s);
    if (prec) launch_prec_x<double, true>(*S.j64, nj, S.s);
    CHK(check_launch());
    for (int j = 0; j < nj; j++) {
        info[j].status = hst[j].status;
        info[j].iterations = hst[j].iter;
        info[j].error = hst[j].err;
        info[j].outer = 0;
        info[j].true_residual = -1.0;
    }
    return TDDA_OK;
}

// FP64 true residual r = E - A x into jo's r vector; rel <- ||r|| / ||E|| for
// the active jobs.
// zero_x: x = 0 first (r = E exactly, since A 0 = 0: no operator
// application, one pass over E).
int true_residual(SolveSet& S, std::vector<double>& rel, const std::vector<int>& active, bool zero_x = false)
{
    const int nj = S.nj;
    SolverState* hst = S.h_st + nj;
    CHK(init_states<double>(S, *S.jo, hst, active, 0.0, 0));
    if (zero_x) launch_init_zero<double>(*S.jo, nj, S.s);
    else launch_apply<double>(*S.jo, nj, IN_X, OUT_RES, 1, 1, S.s);
    CHK(check_launch());
    HIPCHK(hipMemcpyAsync(hst, S.jo->st, sizeof(SolverState) * nj, hipMemcpyDeviceToHost, S.s));
    HIPCHK(hipStreamSynchronize(S.s));
    rel.resize(nj);
    for (int j = 0; j < nj; j++)
        if (active[j]) rel[j] = hst[j].rnorm / hst[j].enorm;
    return TDDA_OK;
}

// Stop test of an inner complex64 solve whose right-hand side is the
// residual r (||r|| / ||E|| = rel): inner_tol, or looser when rel is
// already close to tol -- the correction then only needs to gain the
// remaining factor rel / tol (with a margin of 2).
double inner_tolerance(const tdda_solve_opts* o, double rel)
{
    return std::max(o->inner_tol, 0.5 * o->tol / rel);
}

// One mixed-precision glue launch over the whole set (k_mix).
template <int OP>
void launch_mix(SolveSet& S, int prec)
{
    k_mix<OP><<<dim3(S.j64->nv), dim3(kThreads), 0, S.s>>>(S.j64->d, S.jo->d, S.j32->d, S.j64->tv, S.nj, prec);
}

int solve_mixed(SolveSet& S, const tdda_solve_opts* o, tdda_solve_info* info)
{
    hipStream_t s = S.s;
    const int nj = S.nj;
    if (!S.j32) return fail(TDDA_ERR_STATE, "mixed solve needs operators built with TDDA_C64");
    const int prec = o->precond == TDDA_PRECOND_JACOBI;
    std::vector<int> active(nj, 1);
    std::vector<double> rel(nj, 0.0);
    std::vector<int> inner_it(nj, 0), outer(nj, 0);
    CHK(true_residual(S, rel, active, o->x0_zero != 0));
    SolverState* hin = S.h_st + 2 * nj;
    for (int c = 0; c < o->max_outer; c++) {
        bool any = false;
        for (int j = 0; j < nj; j++) {
            active[j] = active[j] && !(rel[j] <= o->tol) && std::isfinite(rel[j]);
            any |= active[j] != 0;
        }
        if (!any) break;
        std::vector<double> itol(nj);
        for (int j = 0; j < nj; j++) itol[j] = inner_tolerance(o, rel[j]);
        CHK(init_states<float>(S, *S.j32, hin, active, o->inner_tol, o->inner_max_it, &itol, prec));
        // inner right-hand side = r (complex64) of the cycle's jobs
        launch_mix<MIX_DEMOTE>(S, prec);
        launch_init_zero<float>(*S.j32, nj, s);
        CHK(check_launch());
        CHK(run_iterations<float>(S, *S.j32, hin, o->inner_max_it, o->check_every));
        launch_settle<float>(*S.j32, nj, s);
        for (int j = 0; j < nj; j++) {
            if (!active[j]) continue;
            inner_it[j] += hin[j].iter;
            outer[j] += 1;
        }
        launch_mix<MIX_ACCUM>(S, prec);
        CHK(check_launch());
        CHK(true_residual(S, rel, active));
    }
    for (int j = 0; j < nj; j++) {
        info[j].status = (rel[j] <= o->tol) ? TDDA_SOLVE_CONVERGED : TDDA_SOLVE_MAXIT;
        info[j].iterations = inner_it[j];
        info[j].error = rel[j];
        info[j].true_residual = rel[j];
        info[j].outer = outer[j];
    }
    return TDDA_OK;
}

int solve_set(SolveSet& S, const tdda_solve_opts* opts, tdda_solve_info* info)
{
    if (opts->max_it < 0 || !(opts->tol >= 0)) return fail(TDDA_ERR_ARG, "bad iteration limit or tolerance");
    if (opts->precond != TDDA_PRECOND_NONE && opts->precond != TDDA_PRECOND_JACOBI)
        return fail(TDDA_ERR_ARG, "unknown preconditioner");
    if (opts->mode == TDDA_MODE_FP64) {
        CHK(solve_fp64(S, opts, info));
        if (opts->true_residual) {
            std::vector<double> rel(S.nj, 0.0);
            std::vector<int> act(S.nj, 1);
            CHK(true_residual(S, rel, act));
            for (int j = 0; j < S.nj; j++) info[j].true_residual = rel[j];
        }
        return TDDA_OK;
    }
    if (opts->mode == TDDA_MODE_MIXED) {
        if (opts->inner_max_it <= 0 || opts->max_outer <= 0) return fail(TDDA_ERR_ARG, "bad inner/outer limits");
        return solve_mixed(S, opts, info);
    }
    return fail(TDDA_ERR_ARG, "unknown solve mode");
}

}  // namespace

template <typename T>
static int time_passes_t(hipStream_t s, JobTable<T>& jt, int reps, double* ms5)
{
    hipEvent_t ev[6];
    for (int k = 0; k < 6; k++) HIPCHK(hipEventCreate(&ev[k]));
    double acc[5] = {0, 0, 0, 0, 0};
    for (int r = 0; r < reps + 1; r++) {
        HIPCHK(hipEventRecord(ev[0], s));
        run_p1<T>(jt, IN_PLAIN, s);
        HIPCHK(hipEventRecord(ev[1], s));
        run_py<T>(jt, false, 0, s);
        HIPCHK(hipEventRecord(ev[2], s));
        run_px<T>(jt, 0, s);
        HIPCHK(hipEventRecord(ev[3], s));
        run_py<T>(jt, true, 0, s);
        HIPCHK(hipEventRecord(ev[4], s));
        run_p5<T>(jt, OUT_PLAIN, 1, s);
        HIPCHK(hipEventRecord(ev[5], s));
        HIPCHK(hipEventSynchronize(ev[5]));
        if (r == 0) continue;  // warm-up
        for (int k = 0; k < 5; k++) {
            float ms = 0.f;
            HIPCHK(hipEventElapsedTime(&ms, ev[k], ev[k + 1]));
            acc[k] += ms;
        }
    }
    CHK(check_launch());
    for (int k = 0; k < 6; k++) (void)hipEventDestroy(ev[k]);
    for (int k = 0; k < 5; k++) ms5[k] = acc[k] / reps;
    return TDDA_OK;
}

// One pass kind launched reps times back to back between two events (the
// mean is the kernel's duration plus the dispatch gap of a full queue).
template <typename T>
static int time_pass_t(hipStream_t s, JobTable<T>& jt, int pass, int reps, double* ms)
{
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    auto one = [&]() {
        switch (pass) {
        case 0: run_p1<T>(jt, IN_PLAIN, s); break;
        case 1: run_py<T>(jt, false, 0, s); break;
        case 2: run_px<T>(jt, 0, s); break;
        case 3: run_py<T>(jt, true, 0, s); break;
        default: run_p5<T>(jt, OUT_PLAIN, 1, s); break;
        }
    };
    for (int k = 0; k < 3; k++) one();  // warm-up
    HIPCHK(hipEventRecord(e0, s));
    for (int r = 0; r < reps; r++) one();
    HIPCHK(hipEventRecord(e1, s));
    HIPCHK(hipEventSynchronize(e1));
    CHK(check_launch());
    float t = 0.f;
    HIPCHK(hipEventElapsedTime(&t, e0, e1));
    *ms = t / reps;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return TDDA_OK;
}

namespace {
int noalloc_init(int device);  // tdda_noalloc.inc
}

// ---------------------------------------------------------------------------
// exported functions
// ---------------------------------------------------------------------------
extern "C" {

const char* tdda_last_error(void) { return g_err.c_str(); }

int tdda_version(void) { return 100; }

int tdda_ctx_create(int device, tdda_ctx* out)
{
    if (!out) return fail(TDDA_ERR_ARG, "null output handle");
    int n = 0;
    HIPCHK(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) return fail(TDDA_ERR_ARG, "device index out of range");
    HIPCHK(hipSetDevice(device));
    // device address of the allocation-free entry point's static scratch
    // array, looked up once per device with the device current
    CHK(noalloc_init(device));
    hipDeviceProp_t prop;
    HIPCHK(hipGetDeviceProperties(&prop, device));
    tdda_ctx c = new tdda_ctx_s();
    c->device = device;
    c->refs = 1;
    c->ncu = prop.multiProcessorCount;
    // gfx950: 160 KB of LDS per CU, all of it available to one workgroup;
    // only tiles of very long x pencils (c128, Mx > 512) need more than 64 KB
    c->lds_budget = std::min<size_t>(prop.sharedMemPerBlock, 163840);
    c->lds_cu = prop.maxSharedMemoryPerMultiProcessor ? prop.maxSharedMemoryPerMultiProcessor : 65536;
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete c;
        return fail(TDDA_ERR_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(e));
    }
    // auxiliary streams for the launch groups of a pass (default priority);
    // TDDA_AUX_STREAMS=0 keeps every launch on the context's stream
    const int naux = std::max(0, std::min(kAuxStreams, env_int("TDDA_AUX_STREAMS", kAuxStreams)));
    if (naux > 0 && hipEventCreateWithFlags(&c->aux.fork, hipEventDisableTiming) == hipSuccess) {
        for (int k = 0; k < naux; k++) {
            if (hipStreamCreateWithFlags(&c->aux.s[k], hipStreamNonBlocking) != hipSuccess ||
                hipEventCreateWithFlags(&c->aux.join[k], hipEventDisableTiming) != hipSuccess) {
                (void)hipGetLastError();
                if (c->aux.s[k]) (void)hipStreamDestroy(c->aux.s[k]);
                c->aux.s[k] = nullptr;
                break;
            }
            c->aux.n = k + 1;
        }
    }
    (void)hipGetLastError();
    *out = c;
    return TDDA_OK;
}

static void ctx_release(tdda_ctx ctx)
{
    if (--ctx->refs > 0) return;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    for (int k = 0; k < ctx->aux.n; k++) {
        (void)hipStreamSynchronize(ctx->aux.s[k]);
        (void)hipStreamDestroy(ctx->aux.s[k]);
        (void)hipEventDestroy(ctx->aux.join[k]);
    }
    if (ctx->aux.fork) (void)hipEventDestroy(ctx->aux.fork);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

int tdda_ctx_destroy(tdda_ctx ctx)
{
    if (!ctx) return TDDA_OK;
    ctx_release(ctx);
    return TDDA_OK;
}

int tdda_ctx_sync(tdda_ctx ctx)
{
    if (!ctx) return fail(TDDA_ERR_ARG, "null context");
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return TDDA_OK;
}

int tdda_green_tensor_fcd(double K, double d, double x, double y, double z, double* out12)
{
    if (!out12) return fail(TDDA_ERR_ARG, "null output");
    if (!(d > 0) || !(M_PI - K * d > kSiCiStep)) return fail(TDDA_ERR_ARG, "FCD needs d > 0 and K d < pi - 0.1");
    fcd_inter(K, d, x, y, z, out12);
    return TDDA_OK;
}

int tdda_green_tensor(double K, double x, double y, double z, double* out12)
{
    if (!out12) return fail(TDDA_ERR_ARG, "null output");
    const double xs = x * x, ys = y * y, zs = z * z;
    const double rn = std::sqrt(xs + ys + zs);
    if (rn == 0.0) {
        for (int l = 0; l < 12; l++) out12[l] = 0.0;
        return TDDA_OK;
    }
    const double r2 = rn * rn, r3 = rn * rn * rn;
    const double c1r = std::cos(K * rn) / r3, c1i = std::sin(K * rn) / r3;
    const double c2r = 1.0 / r2, c2i = -K * rn / r2;
    const double c3r = K * K - 3.0 * c2r, c3i = -3.0 * c2i;
    const double kk = K * K;
    double a[6][2] = {{c2r * (r2 - 3.0 * xs) - kk * (ys + zs), c2i * (r2 - 3.0 * xs)},
                      {c3r * x * y, c3i * x * y},
                      {c3r * z * x, c3i * z * x},
                      {c2r * (r2 - 3.0 * ys) - kk * (zs + xs), c2i * (r2 - 3.0 * ys)},
                      {c3r * y * z, c3i * y * z},
                      {c2r * (r2 - 3.0 * zs) - kk * (xs + ys), c2i * (r2 - 3.0 * zs)}};
    for (int l = 0; l < 6; l++) {
        out12[2 * l] = c1r * a[l][0] - c1i * a[l][1];
        out12[2 * l + 1] = c1r * a[l][1] + c1i * a[l][0];
    }
    return TDDA_OK;
}

// |eps| as the reference computes it, sqrt(norm(eps)), so that the 1.01
// threshold falls on the same side for every input.
static int material_rule(int nmat, double* material, double nback, bool strict)
{
    if (nmat < 0 || (nmat > 0 && !material) || !(nback > 0)) return fail(TDDA_ERR_ARG, "bad material list");
    for (int i = 0; i < nmat; i++) {
        std::complex<double> m(material[2 * i], material[2 * i + 1]);
        m = m / (nback * nback);
        const double a = std::sqrt(std::norm(m));
        if (strict ? a < 1.01 : a <= 1.01) m = m + 0.01;
        material[2 * i] = m.real();
        material[2 * i + 1] = m.imag();
    }
    return TDDA_OK;
}

int tdda_material_rule(int nmat, double* material, double nback) { return material_rule(nmat, material, nback, false); }

int tdda_material_rule_periodic(int nmat, double* material, double nback)
{
    return material_rule(nmat, material, nback, true);
}

static std::complex<double> ldr_alpha(double K, double d, std::complex<double> diel, const double* nE,
                                      const double* nK)
{
    const double b1 = -1.891531, b2 = 0.1648469, b3 = -1.7700004;
    const std::complex<double> a1 = (3.0 * std::pow(d, 3) / (4.0 * M_PI)) * (diel - 1.0) / (diel + 2.0);
    std::complex<double> res(0.0, (2.0 / 3.0) * std::pow(K * d, 3));
    const double S = std::pow(nE[0] * nK[0], 2) + std::pow(nE[1] * nK[1], 2) + std::pow(nE[2] * nK[2], 2);
    res = 1.0 + (a1 / std::pow(d, 3)) * ((b1 + diel * b2 + diel * b3 * S) * std::pow(K * d, 2) - res);
    return a1 / res;
}

int tdda_get_alpha(double lam, double K, double d, double diel_re, double diel_im, const double* nE3,
                   const double* nK3, double* out2)
{
    (void)lam;
    if (!nE3 || !nK3 || !out2) return fail(TDDA_ERR_ARG, "null argument");
    const std::complex<double> a = ldr_alpha(K, d, std::complex<double>(diel_re, diel_im), nE3, nK3);
    out2[0] = a.real();
    out2[1] = a.imag();
    return TDDA_OK;
}

int tdda_incident_field(int N, const int* R, double K, double d, double E0, const double* nK3, const double* nE3,
                        double* E)
{
    if (N < 0 || (N > 0 && (!R || !E)) || !nK3 || !nE3) return fail(TDDA_ERR_ARG, "null argument");
    for (int i = 0; i < N; i++) {
        const double ph = K * d * (nK3[0] * R[3 * i] + nK3[1] * R[3 * i + 1] + nK3[2] * R[3 * i + 2]);
        const double c = std::cos(ph), sn = std::sin(ph);
        for (int k = 0; k < 3; k++) {
            E[2 * (3 * i + k)] = E0 * nE3[k] * c;
            E[2 * (3 * i + k) + 1] = E0 * nE3[k] * sn;
        }
    }
    return TDDA_OK;
}

int tdda_polarizability(int n3, const double* diel_old, int nmat, const double* material, double lam, double K,
                        double d, const double* nE3, const double* nK3, double* diel, double* al)
{
    (void)lam;
    if (n3 < 0 || (n3 > 0 && (!diel_old || !al)) || nmat <= 0 || !material || !nE3 || !nK3)
        return fail(TDDA_ERR_ARG, "bad argument");
    for (int i = 0; i < n3; i++) {
        const double t = diel_old[i];
        const int lf = (int)std::floor(t);
        int ln = lf + 1;
        if (lf >= 1) ln = lf;
        if (lf < 0 || ln >= nmat) return fail(TDDA_ERR_ARG, "design value outside the material table");
        const std::complex<double> mf(material[2 * lf], material[2 * lf + 1]);
        const std::complex<double> mn(material[2 * ln], material[2 * ln + 1]);
        const std::complex<double> e = mf + (t - (double)lf) * (mn - mf);
        const std::complex<double> a = 1.0 / ldr_alpha(K, d, e, nE3, nK3);
        if (diel) {
            diel[2 * i] = e.real();
            diel[2 * i + 1] = e.imag();
        }
        al[2 * i] = a.real();
        al[2 * i + 1] = a.imag();
    }
    return TDDA_OK;
}

int tdda_local_field(int n3, const double* al, const double* P, const double* diel, double* out)
{
    if (n3 < 0 || (n3 > 0 && (!al || !P || !diel || !out))) return fail(TDDA_ERR_ARG, "null argument");
    for (int i = 0; i < n3; i++) {
        const std::complex<double> lf = (2.0 + std::complex<double>(diel[2 * i], diel[2 * i + 1])) / 3.0;
        const std::complex<double> v =
            std::complex<double>(al[2 * i], al[2 * i + 1]) * std::complex<double>(P[2 * i], P[2 * i + 1]) / lf;
        out[2 * i] = v.real();
        out[2 * i + 1] = v.imag();
    }
    return TDDA_OK;
}

namespace {

// Images of a structure periodic in x and y (pm < 0: not periodic) and the
// interaction tensor (0: LDR, 1: FCD).
struct PeriodicImages {
    int pm = -1, pn = -1;
    double Lm = 0.0, Ln = 0.0;
    int method = TDDA_METHOD_LDR;
};

int op_create(tdda_ctx ctx, const tdda_op_desc* desc, const PeriodicImages& per, tdda_op* out);

}  // namespace

int tdda_op_create(tdda_ctx ctx, const tdda_op_desc* desc, tdda_op* out)
{
    return op_create(ctx, desc, PeriodicImages(), out);
}

int tdda_op_create_ex(tdda_ctx ctx, const tdda_op_desc* desc, const tdda_op_ext* ext, tdda_op* out)
{
    if (!ext) return fail(TDDA_ERR_ARG, "null argument");
    if (ext->method != TDDA_METHOD_LDR && ext->method != TDDA_METHOD_FCD)
        return fail(TDDA_ERR_ARG, "unknown interaction method");
    PeriodicImages per;
    if (ext->max_m >= 0 || ext->max_n >= 0) {
        if (ext->max_m < 0 || ext->max_n < 0) return fail(TDDA_ERR_ARG, "image counts must both be >= 0");
        if (!(ext->Lm > 0) || !(ext->Ln > 0) || !std::isfinite(ext->Lm) || !std::isfinite(ext->Ln))
            return fail(TDDA_ERR_ARG, "periods must be positive");
        if ((double)(2 * ext->max_m + 1) * (2 * ext->max_n + 1) > 1e6)
            return fail(TDDA_ERR_UNSUPPORTED, "more than 1e6 images");
        per.pm = ext->max_m;
        per.pn = ext->max_n;
        per.Lm = ext->Lm;
        per.Ln = ext->Ln;
    }
    per.method = ext->method;
    return op_create(ctx, desc, per, out);
}

int tdda_op_create_periodic(tdda_ctx ctx, const tdda_op_desc* desc, int max_m, int max_n, double Lm, double Ln,
                            tdda_op* out)
{
    if (max_m < 0 || max_n < 0) return fail(TDDA_ERR_ARG, "image counts must be >= 0");
    if (!(Lm > 0) || !(Ln > 0) || !std::isfinite(Lm) || !std::isfinite(Ln))
        return fail(TDDA_ERR_ARG, "periods must be positive");
    if ((double)(2 * max_m + 1) * (2 * max_n + 1) > 1e6) return fail(TDDA_ERR_UNSUPPORTED, "more than 1e6 images");
    PeriodicImages per;
    per.pm = max_m;
    per.pn = max_n;
    per.Lm = Lm;
    per.Ln = Ln;
    return op_create(ctx, desc, per, out);
}

}  // extern "C"

namespace {

// Checks of an operator description that need no device; M <- the FFT grid.
int check_desc(const tdda_op_desc& D, int* M)
{
    if (D.N <= 0 || D.Nx <= 0 || D.Ny <= 0 || D.Nz <= 0) return fail(TDDA_ERR_ARG, "lattice sizes must be positive");
    if (!(D.d > 0) || !(D.lam > 0) || !(D.nback > 0)) return fail(TDDA_ERR_ARG, "d, lam and nback must be positive");
    const long long box = (long long)D.Nx * D.Ny * D.Nz;
    if (box > (1LL << 30) || D.N > box) return fail(TDDA_ERR_ARG, "lattice too large or N exceeds the box");
    const int prec = D.precisions ? D.precisions : (TDDA_C128 | TDDA_C64);
    if (prec & ~(TDDA_C128 | TDDA_C64)) return fail(TDDA_ERR_ARG, "unknown precision bits");
    M[0] = D.Mx;
    M[1] = D.My;
    M[2] = D.Mz;
    const int Nn[3] = {D.Nx, D.Ny, D.Nz};
    for (int a = 0; a < 3; a++) {
        if (M[a] == 0) M[a] = default_grid_len(a, 2 * Nn[a] - 1, prec);
        if (M[a] < 2 * Nn[a] - 1) return fail(TDDA_ERR_ARG, "FFT grid smaller than 2N-1 along an axis");
        if (!smooth235(M[a])) return fail(TDDA_ERR_UNSUPPORTED, "FFT grid sizes must have prime factors 2, 3, 5 only");
        if (M[a] > 4096) return fail(TDDA_ERR_UNSUPPORTED, "FFT length above 4096");
    }
    // the pass kernels address one right-hand side's intermediates (the
    // largest, after the y pass, holds Nx x My x Mz x 3 values; the site
    // vectors 3N are smaller) with 32-bit offsets
    if ((long long)D.Nx * M[1] * M[2] * 3 >= (1LL << 31))
        return fail(TDDA_ERR_UNSUPPORTED, "structure too large: Nx * My * Mz * 3 must stay below 2^31");
    return TDDA_OK;
}

int op_create(tdda_ctx ctx, const tdda_op_desc* desc, const PeriodicImages& per, tdda_op* out)
{
    if (!ctx || !desc || !out) return fail(TDDA_ERR_ARG, "null argument");
    const tdda_op_desc& D = *desc;
    int M[3];
    CHK(check_desc(D, M));
    const long long box = (long long)D.Nx * D.Ny * D.Nz;
    const int prec = D.precisions ? D.precisions : (TDDA_C128 | TDDA_C64);
    HIPCHK(hipSetDevice(ctx->device));
    // bulk order (tools.cpp:497-510) on the full box needs no cell map;
    // R == NULL stands for that order
    if (!D.R && box != D.N) return fail(TDDA_ERR_ARG, "R may be NULL only for a full box");
    bool bulk = (box == D.N);
    std::vector<int> cmap(D.R ? (size_t)box : 0, -1);
    for (int i = 0; D.R && i < D.N; i++) {
        const int x = D.R[3 * i], y = D.R[3 * i + 1], z = D.R[3 * i + 2];
        if (x < 0 || x >= D.Nx || y < 0 || y >= D.Ny || z < 0 || z >= D.Nz) {
            return fail(TDDA_ERR_ARG, "site coordinate outside the Nx x Ny x Nz box");
        }
        const size_t cell = ((size_t)x * D.Ny + y) * D.Nz + z;
        if (cmap[cell] != -1) {
            return fail(TDDA_ERR_ARG, "two sites share one lattice cell");
        }
        cmap[cell] = i;
        if ((size_t)i != cell) bulk = false;
    }
    tdda_op op = new tdda_op_s();
    op->ctx = ctx;
    ctx->refs++;
    op->N = D.N; op->Nx = D.Nx; op->Ny = D.Ny; op->Nz = D.Nz;
    op->Mx = M[0]; op->My = M[1]; op->Mz = M[2];
    op->Hx = M[0] / 2 + 1; op->Hy = M[1] / 2 + 1; op->Hz = M[2] / 2 + 1;
    op->d = D.d; op->lam = D.lam; op->nback = D.nback;
    op->K = 2.0 * M_PI / (D.lam / D.nback);  // AProductCore.cpp:17
    if (per.method == TDDA_METHOD_FCD && !(M_PI - op->K * D.d > kSiCiStep)) {
        // the smallest Si / Ci argument, (pi / d - K) d, must lie inside the
        // reference's tables (x >= 0.1)
        tdda_op_destroy(op);
        return fail(TDDA_ERR_UNSUPPORTED, "FCD needs K d < pi - 0.1");
    }
    op->prec = prec;
    op->method = per.method;
    auto cleanup = [&](int rc) {
        tdda_op_destroy(op);
        return rc;
    };
    if (!bulk) {
        if (ctx_malloc(ctx, (void**)&op->cellmap, sizeof(int) * box) != TDDA_OK) return cleanup(fail(TDDA_ERR_OOM, "cell map"));
        if (hipMemcpy(op->cellmap, cmap.data(), sizeof(int) * box, hipMemcpyHostToDevice) != hipSuccess)
            return cleanup(fail(TDDA_ERR_HIP, "cell map upload"));
    }
    for (int a = 0; a < 3; a++) {
        op->pl[a] = make_plan(M[a]);
        std::vector<Cx<double>> t = twiddles(M[a]);
        std::vector<Cx<float>> tf(M[a]);
        for (int q = 0; q < M[a]; q++) tf[q] = Cx<float>{(float)t[q].x, (float)t[q].y};
        if (ctx_malloc(ctx, (void**)&op->tw64[a], sizeof(Cx<double>) * M[a]) != TDDA_OK ||
            ctx_malloc(ctx, (void**)&op->tw32[a], sizeof(Cx<float>) * M[a]) != TDDA_OK)
            return cleanup(fail(TDDA_ERR_OOM, "twiddles"));
        if (hipMemcpy(op->tw64[a], t.data(), sizeof(Cx<double>) * M[a], hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(op->tw32[a], tf.data(), sizeof(Cx<float>) * M[a], hipMemcpyHostToDevice) != hipSuccess)
            return cleanup(fail(TDDA_ERR_HIP, "twiddle upload"));
        std::vector<Cx<double>> t2 = twiddles2(M[a], op->pl[a].tb);
        std::vector<Cx<float>> t2f(t2.size());
        for (size_t q = 0; q < t2.size(); q++) t2f[q] = Cx<float>{(float)t2[q].x, (float)t2[q].y};
        if (ctx_malloc(ctx, (void**)&op->tws64[a], sizeof(Cx<double>) * t2.size()) != TDDA_OK ||
            ctx_malloc(ctx, (void**)&op->tws32[a], sizeof(Cx<float>) * t2.size()) != TDDA_OK)
            return cleanup(fail(TDDA_ERR_OOM, "twiddles"));
        if (hipMemcpy(op->tws64[a], t2.data(), sizeof(Cx<double>) * t2.size(), hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(op->tws32[a], t2f.data(), sizeof(Cx<float>) * t2.size(), hipMemcpyHostToDevice) != hipSuccess)
            return cleanup(fail(TDDA_ERR_HIP, "twiddle upload"));
    }
    // Green tensor on the full grid, FP64, then its 3D forward transform
    // (AProductCore.cpp:54-229).
    const size_t Mt = (size_t)M[0] * M[1] * M[2];
    Cx<double>* full = nullptr;
    if (ctx_malloc(ctx, (void**)&full, sizeof(Cx<double>) * Mt * 6) != TDDA_OK)
        return cleanup(fail(TDDA_ERR_OOM, "Green tensor grid"));
    hipStream_t s = ctx->stream;
    k_green_fill<<<(unsigned)((Mt + kThreads - 1) / kThreads), kThreads, 0, s>>>(
        full, D.Nx, D.Ny, D.Nz, M[0], M[1], M[2], D.d, op->K, per.pm, per.pn, per.Lm, per.Ln,
        per.method == TDDA_METHOD_FCD ? 1 : 0);
    {
        // z: contiguous pencils [x][y][z][6]
        const size_t budget = ctx->lds_budget;
        const int ps = M[2] + 1;
        int P = (int)(budget / (2 * 6 * ps * sizeof(Cx<double>)));
        P = std::max(1, std::min(P, 32));
        if (2 * 6 * (size_t)ps * sizeof(Cx<double>) > 65536) {
            ctx_free(ctx, full);
            return cleanup(fail(TDDA_ERR_UNSUPPORTED, "z FFT length too large for the Green set-up pass"));
        }
        const int Q = M[0] * M[1];
        k_fft_contig<double, false><<<(Q + P - 1) / P, kThreads, 2 * 6 * P * ps * sizeof(Cx<double>), s>>>(
            full, Q, M[2], 6, P, ps, op->pl[2], op->tw64[2]);
        // y: outer x, element (x, y, f) at x*My*Mz*6 + y*Mz*6 + f
        const int F = M[2] * 6;
        for (int ax = 1; ax >= 0; ax--) {
            const int L = M[ax];
            const int psa = L + 1;
            int tf = (int)(budget / (2 * psa * sizeof(Cx<double>)));
            tf = std::max(1, std::min(tf, 48));
            const int ntf = (F + tf - 1) / tf;
            const int outer = (ax == 1) ? M[0] : M[1];
            const size_t so = (ax == 1) ? (size_t)M[1] * F : (size_t)F;
            const size_t sa = (ax == 1) ? (size_t)F : (size_t)M[1] * F;
            k_fft_strided<double, false><<<outer * ntf, kThreads, 2 * tf * psa * sizeof(Cx<double>), s>>>(
                full, L, so, sa, F, tf, ntf, psa, op->pl[ax], op->tw64[ax]);
        }
    }
    const size_t Ht = (size_t)op->Hx * op->Hy * op->Hz * 6;
    const double scale = 1.0 / (double)Mt;
    const unsigned gb = (unsigned)((Ht + kThreads - 1) / kThreads);
    if (prec & TDDA_C128) {
        if (ctx_malloc(ctx, (void**)&op->green64, sizeof(Cx<double>) * Ht) != TDDA_OK) {
            ctx_free(ctx, full);
            return cleanup(fail(TDDA_ERR_OOM, "Green spectrum"));
        }
        k_green_octant<double><<<gb, kThreads, 0, s>>>(full, op->green64, M[0], M[1], M[2], op->Hx, op->Hy, op->Hz,
                                                       scale);
    }
    if (prec & TDDA_C64) {
        if (ctx_malloc(ctx, (void**)&op->green32, sizeof(Cx<float>) * Ht) != TDDA_OK) {
            ctx_free(ctx, full);
            return cleanup(fail(TDDA_ERR_OOM, "Green spectrum"));
        }
        k_green_octant<float><<<gb, kThreads, 0, s>>>(full, op->green32, M[0], M[1], M[2], op->Hx, op->Hy, op->Hz,
                                                      scale);
    }
    hipError_t e = hipGetLastError();
    hipError_t e2 = hipStreamSynchronize(s);
    ctx_free(ctx, full);
    if (e != hipSuccess || e2 != hipSuccess)
        return cleanup(fail(TDDA_ERR_HIP, std::string("Green set-up: ") + hipGetErrorString(e != hipSuccess ? e : e2)));
    {
        int rc = TDDA_OK;
        if ((prec & TDDA_C128) && (rc = plan_tiles<double>(op, op->tp64)) != TDDA_OK) return cleanup(rc);
        if ((prec & TDDA_C64) && (rc = plan_tiles<float>(op, op->tp32)) != TDDA_OK) return cleanup(rc);
    }
    *out = op;
    return TDDA_OK;
}

}  // namespace

extern "C" {

static void op_release(tdda_op op)
{
    if (--op->refs > 0) return;
    (void)hipSetDevice(op->ctx->device);
    (void)hipStreamSynchronize(op->ctx->stream);
    ctx_free(op->ctx, op->cellmap);
    ctx_free(op->ctx, op->site_xyz);
    ctx_free(op->ctx, op->green64);
    ctx_free(op->ctx, op->green32);
    for (int a = 0; a < 3; a++) {
        ctx_free(op->ctx, op->tw64[a]);
        ctx_free(op->ctx, op->tw32[a]);
        ctx_free(op->ctx, op->tws64[a]);
        ctx_free(op->ctx, op->tws32[a]);
    }
    tdda_ctx ctx = op->ctx;
    delete op;
    ctx_release(ctx);
}

int tdda_op_destroy(tdda_op op)
{
    if (!op) return TDDA_OK;
    if (op->scratch) {
        tdda_model m = op->scratch;
        op->scratch = nullptr;
        tdda_model_destroy(m);
    }
    op_release(op);
    return TDDA_OK;
}

int tdda_op_plan(const tdda_op_desc* desc, int* grid3)
{
    if (!desc) return fail(TDDA_ERR_ARG, "null argument");
    int M[3];
    CHK(check_desc(*desc, M));
    if (grid3)
        for (int a = 0; a < 3; a++) grid3[a] = M[a];
    return TDDA_OK;
}

int tdda_op_info(tdda_op op, int* grid3, double* K)
{
    if (!op) return fail(TDDA_ERR_ARG, "null operator");
    if (grid3) {
        grid3[0] = op->Mx;
        grid3[1] = op->My;
        grid3[2] = op->Mz;
    }
    if (K) *K = op->K;
    return TDDA_OK;
}

int tdda_model_create(tdda_op op, int nrhs, tdda_model* out)
{
    if (!op || !out) return fail(TDDA_ERR_ARG, "null argument");
    if (nrhs <= 0 || nrhs > 4096) return fail(TDDA_ERR_ARG, "nrhs out of range");
    HIPCHK(hipSetDevice(op->ctx->device));
    tdda_model m = new tdda_model_s();
    m->op = op;
    op->refs++;
    m->nrhs = nrhs;
    m->n3 = 3 * (size_t)op->N;
    auto bail = [&](int rc) {
        tdda_model_destroy(m);
        return rc;
    };
    int rc;
    if (op->green64) {
        if ((rc = alloc_vecs<double>(m, m->v64)) != TDDA_OK) return bail(rc);
        Cx<double>* tw[6] = {op->tws64[0], op->tws64[1], op->tws64[2], op->tw64[0], op->tw64[1], op->tw64[2]};
        fill_jobs<double>(op, nrhs, m->j64.h, m->v64, op->green64, tw, op->tp64, m->v64.r, m->v64.r0);
        // outer residual of the mixed solver: r -> p, its r0 copy -> t (scratch)
        fill_jobs<double>(op, nrhs, m->jo.h, m->v64, op->green64, tw, op->tp64, m->v64.p, m->v64.t);
        if ((rc = finalize_table<double>(m->allocs, op->ctx, m->j64)) != TDDA_OK) return bail(rc);
        if ((rc = finalize_table<double>(m->allocs, op->ctx, m->jo)) != TDDA_OK) return bail(rc);
    }
    if (op->green32) {
        if ((rc = alloc_vecs<float>(m, m->v32)) != TDDA_OK) return bail(rc);
        Cx<float>* tw[6] = {op->tws32[0], op->tws32[1], op->tws32[2], op->tw32[0], op->tw32[1], op->tw32[2]};
        fill_jobs<float>(op, nrhs, m->j32.h, m->v32, op->green32, tw, op->tp32, m->v32.r, m->v32.r0);
        if ((rc = finalize_table<float>(m->allocs, op->ctx, m->j32)) != TDDA_OK) return bail(rc);
    }
    if (hipStreamSynchronize(op->ctx->stream) != hipSuccess) return bail(fail(TDDA_ERR_HIP, "model set-up"));
    *out = m;
    return TDDA_OK;
}

static void model_release(tdda_model m);

int tdda_model_destroy(tdda_model m)
{
    if (!m) return TDDA_OK;
    model_release(m);
    return TDDA_OK;
}

static void model_release(tdda_model m)
{
    if (--m->refs > 0) return;
    (void)hipSetDevice(m->op->ctx->device);
    (void)hipStreamSynchronize(m->op->ctx->stream);
    for (void* p : m->allocs) ctx_free(m->op->ctx, p);
    if (m->h_st) (void)hipHostFree(m->h_st);
    if (m->h_poll) (void)hipHostFree(m->h_poll);
    for (int k = 0; k < 2; k++)
        if (m->ev[k]) (void)hipEventDestroy(m->ev[k]);
    ctx_free(m->op->ctx, m->d_idx);
    ctx_free(m->op->ctx, m->d_vals);
    if (m->h_idx) (void)hipHostFree(m->h_idx);
    if (m->h_vals) (void)hipHostFree(m->h_vals);
    tdda_op op = m->op;
    delete m;
    op_release(op);
}

static Cx<double>* vec64(tdda_model m, int which)
{
    switch (which) {
    case TDDA_VEC_AL: return m->v64.al;
    case TDDA_VEC_E: return m->v64.e;
    case TDDA_VEC_P: return m->v64.x;
    case TDDA_VEC_Y: return m->v64.ap;
    default: return nullptr;
    }
}

int tdda_model_set(tdda_model m, int which, int rhs, const double* host)
{
    if (!m || !host) return fail(TDDA_ERR_ARG, "null argument");
    if (rhs < -1 || rhs >= m->nrhs) return fail(TDDA_ERR_ARG, "rhs index out of range");
    if (which != TDDA_VEC_AL && which != TDDA_VEC_E && which != TDDA_VEC_P)
        return fail(TDDA_ERR_ARG, "vector kind not settable");
    if (!m->v64.x) return fail(TDDA_ERR_STATE, "model has no complex128 storage");
    HIPCHK(hipSetDevice(m->op->ctx->device));
    hipStream_t s = m->op->ctx->stream;
    const size_t n3 = m->n3;
    const size_t off = rhs < 0 ? 0 : (size_t)rhs * n3;
    const size_t cnt = rhs < 0 ? n3 * m->nrhs : n3;
    Cx<double>* dst = vec64(m, which) + off;
    HIPCHK(hipMemcpyAsync(dst, host, cnt * sizeof(Cx<double>), hipMemcpyHostToDevice, s));
    if (which == TDDA_VEC_AL && m->v32.al) {
        const unsigned blocks = (unsigned)((cnt + kThreads - 1) / kThreads);
        k_demote<<<blocks, kThreads, 0, s>>>(dst, m->v32.al + off, cnt);
        CHK(check_launch());
    }
    HIPCHK(hipStreamSynchronize(s));
    return TDDA_OK;
}

int tdda_model_get(tdda_model m, int which, int rhs, double* host)
{
    if (!m || !host) return fail(TDDA_ERR_ARG, "null argument");
    if (rhs < -1 || rhs >= m->nrhs) return fail(TDDA_ERR_ARG, "rhs index out of range");
    Cx<double>* src = vec64(m, which);
    if (!src) return fail(TDDA_ERR_ARG, "unknown vector kind");
    HIPCHK(hipSetDevice(m->op->ctx->device));
    hipStream_t s = m->op->ctx->stream;
    const size_t n3 = m->n3;
    const size_t off = rhs < 0 ? 0 : (size_t)rhs * n3;
    const size_t cnt = rhs < 0 ? n3 * m->nrhs : n3;
    HIPCHK(hipMemcpyAsync(host, src + off, cnt * sizeof(Cx<double>), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    return TDDA_OK;
}

static int ensure_entries(tdda_model m, int n)
{
    if (n <= m->cap_entries) return TDDA_OK;
    int cap = std::max(n, 2 * m->cap_entries);
    ctx_free(m->op->ctx, m->d_idx);
    ctx_free(m->op->ctx, m->d_vals);
    if (m->h_idx) (void)hipHostFree(m->h_idx);
    if (m->h_vals) (void)hipHostFree(m->h_vals);
    m->d_idx = nullptr;
    m->d_vals = nullptr;
    m->h_idx = nullptr;
    m->h_vals = nullptr;
    m->cap_entries = 0;
    if (ctx_malloc(m->op->ctx, (void**)&m->d_idx, sizeof(int) * cap) != TDDA_OK ||
        ctx_malloc(m->op->ctx, (void**)&m->d_vals, sizeof(Cx<double>) * cap) != TDDA_OK)
        return fail(TDDA_ERR_OOM, "entry staging");
    HIPCHK(hipHostMalloc((void**)&m->h_idx, sizeof(int) * cap, hipHostMallocDefault));
    HIPCHK(hipHostMalloc((void**)&m->h_vals, sizeof(Cx<double>) * cap, hipHostMallocDefault));
    m->cap_entries = cap;
    return TDDA_OK;
}

extern "C" int tdda_model_get_entries(tdda_model m, int which, int rhs, int n, const int* idx, double* vals)
{
    if (!m || (n > 0 && (!idx || !vals))) return fail(TDDA_ERR_ARG, "null argument");
    if (rhs < 0 || rhs >= m->nrhs || n < 0) return fail(TDDA_ERR_ARG, "rhs index or count out of range");
    Cx<double>* src = vec64(m, which);
    if (!src) return fail(TDDA_ERR_ARG, "unknown vector kind");
    for (int k = 0; k < n; k++)
        if (idx[k] < 0 || (size_t)idx[k] >= m->n3) return fail(TDDA_ERR_ARG, "entry index out of range");
    if (n == 0) return TDDA_OK;
    HIPCHK(hipSetDevice(m->op->ctx->device));
    CHK(ensure_entries(m, n));
    hipStream_t s = m->op->ctx->stream;
    memcpy(m->h_idx, idx, sizeof(int) * n);
    HIPCHK(hipMemcpyAsync(m->d_idx, m->h_idx, sizeof(int) * n, hipMemcpyHostToDevice, s));
    k_gather<<<(n + kThreads - 1) / kThreads, kThreads, 0, s>>>(src + (size_t)rhs * m->n3, m->d_idx, n, m->d_vals);
    CHK(check_launch());
    HIPCHK(hipMemcpyAsync(m->h_vals, m->d_vals, sizeof(Cx<double>) * n, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    memcpy(vals, m->h_vals, sizeof(Cx<double>) * n);
    return TDDA_OK;
}

extern "C" int tdda_model_set_entries(tdda_model m, int which, int rhs, int n, const int* idx, const double* vals)
{
    if (!m || (n > 0 && (!idx || !vals))) return fail(TDDA_ERR_ARG, "null argument");
    if (rhs < 0 || rhs >= m->nrhs || n < 0) return fail(TDDA_ERR_ARG, "rhs index or count out of range");
    if (which != TDDA_VEC_AL && which != TDDA_VEC_E && which != TDDA_VEC_P)
        return fail(TDDA_ERR_ARG, "vector kind not settable");
    for (int k = 0; k < n; k++)
        if (idx[k] < 0 || (size_t)idx[k] >= m->n3) return fail(TDDA_ERR_ARG, "entry index out of range");
    HIPCHK(hipSetDevice(m->op->ctx->device));
    CHK(ensure_entries(m, std::max(n, 1)));
    hipStream_t s = m->op->ctx->stream;
    Cx<double>* dst = vec64(m, which) + (size_t)rhs * m->n3;
    HIPCHK(hipMemsetAsync(dst, 0, m->n3 * sizeof(Cx<double>), s));
    if (n > 0) {
        memcpy(m->h_idx, idx, sizeof(int) * n);
        memcpy(m->h_vals, vals, sizeof(Cx<double>) * n);
        HIPCHK(hipMemcpyAsync(m->d_idx, m->h_idx, sizeof(int) * n, hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(m->d_vals, m->h_vals, sizeof(Cx<double>) * n, hipMemcpyHostToDevice, s));
        k_scatter<<<(n + kThreads - 1) / kThreads, kThreads, 0, s>>>(dst, m->d_idx, n, m->d_vals);
        CHK(check_launch());
    }
    if (which == TDDA_VEC_AL && m->v32.al) {
        const unsigned blocks = (unsigned)((m->n3 + kThreads - 1) / kThreads);
        k_demote<<<blocks, kThreads, 0, s>>>(dst, m->v32.al + (size_t)rhs * m->n3, m->n3);
        CHK(check_launch());
    }
    HIPCHK(hipStreamSynchronize(s));
    return TDDA_OK;
}


int tdda_model_time_passes(tdda_model m, int precision, int reps, double* ms5)
{
    if (!m || !ms5 || reps <= 0) return fail(TDDA_ERR_ARG, "bad argument");
    HIPCHK(hipSetDevice(m->op->ctx->device));
    if (precision == TDDA_C128) {
        if (!m->v64.x) return fail(TDDA_ERR_STATE, "operator built without TDDA_C128");
        return time_passes_t<double>(m->op->ctx->stream, m->j64, reps, ms5);
    }
    if (precision == TDDA_C64) {
        if (!m->v32.x) return fail(TDDA_ERR_STATE, "operator built without TDDA_C64");
        const size_t n = m->n3 * m->nrhs;
        k_demote<<<(unsigned)((n + kThreads - 1) / kThreads), kThreads, 0, m->op->ctx->stream>>>(m->v64.x, m->v32.x, n);
        return time_passes_t<float>(m->op->ctx->stream, m->j32, reps, ms5);
    }
    return fail(TDDA_ERR_ARG, "precision must be TDDA_C128 or TDDA_C64");
}

int tdda_model_time_pass(tdda_model m, int precision, int pass, int reps, double* ms)
{
    if (!m || !ms || reps <= 0 || pass < 0 || pass > 4) return fail(TDDA_ERR_ARG, "bad argument");
    HIPCHK(hipSetDevice(m->op->ctx->device));
    if (precision == TDDA_C128) {
        if (!m->v64.x) return fail(TDDA_ERR_STATE, "operator built without TDDA_C128");
        return time_pass_t<double>(m->op->ctx->stream, m->j64, pass, reps, ms);
    }
    if (precision == TDDA_C64) {
        if (!m->v32.x) return fail(TDDA_ERR_STATE, "operator built without TDDA_C64");
        const size_t n = m->n3 * m->nrhs;
        k_demote<<<(unsigned)((n + kThreads - 1) / kThreads), kThreads, 0, m->op->ctx->stream>>>(m->v64.x, m->v32.x, n);
        return time_pass_t<float>(m->op->ctx->stream, m->j32, pass, reps, ms);
    }
    return fail(TDDA_ERR_ARG, "precision must be TDDA_C128 or TDDA_C64");
}

static int model_apply(tdda_model m, int precision, int diag)
{
    hipStream_t s = m->op->ctx->stream;
    const int nj = m->nrhs;
    const size_t n = m->n3 * nj;
    if (precision == TDDA_C128) {
        if (!m->v64.x) return fail(TDDA_ERR_STATE, "operator built without TDDA_C128");
        launch_apply<double>(m->j64, nj, IN_PLAIN, OUT_PLAIN, diag, 0, s);
    } else if (precision == TDDA_C64) {
        if (!m->v32.x) return fail(TDDA_ERR_STATE, "operator built without TDDA_C64");
        const unsigned blocks = (unsigned)((n + kThreads - 1) / kThreads);
        k_demote<<<blocks, kThreads, 0, s>>>(m->v64.x, m->v32.x, n);
        launch_apply<float>(m->j32, nj, IN_PLAIN, OUT_PLAIN, diag, 0, s);
        // promote: y64 = c128(y32)
        HIPCHK(hipMemsetAsync(m->v64.ap, 0, n * sizeof(Cx<double>), s));
        k_accum<<<blocks, kThreads, 0, s>>>(m->v64.ap, m->v32.ap, n);
    } else {
        return fail(TDDA_ERR_ARG, "precision must be TDDA_C128 or TDDA_C64");
    }
    CHK(check_launch());
    HIPCHK(hipStreamSynchronize(s));
    return TDDA_OK;
}

int tdda_model_apply(tdda_model m, int precision)
{
    if (!m) return fail(TDDA_ERR_ARG, "null model");
    HIPCHK(hipSetDevice(m->op->ctx->device));
    return model_apply(m, precision, 1);
}

int tdda_aproduct(tdda_op op, const double* b, double* y, int nrhs, const double* al, int al_per_rhs, int precision)
{
    if (!op || !b || !y) return fail(TDDA_ERR_ARG, "null argument");
    if (nrhs <= 0) return fail(TDDA_ERR_ARG, "nrhs must be positive");
    if (!op->green64) return fail(TDDA_ERR_STATE, "host-buffer Aproduct needs an operator built with TDDA_C128");
    HIPCHK(hipSetDevice(op->ctx->device));
    if (op->scratch && op->scratch->nrhs != nrhs) {
        tdda_model_destroy(op->scratch);
        op->scratch = nullptr;
    }
    if (!op->scratch) CHK(tdda_model_create(op, nrhs, &op->scratch));
    tdda_model m = op->scratch;
    const size_t n3 = m->n3;
    CHK(tdda_model_set(m, TDDA_VEC_P, -1, b));
    if (al) {
        if (al_per_rhs) {
            CHK(tdda_model_set(m, TDDA_VEC_AL, -1, al));
        } else {
            for (int j = 0; j < nrhs; j++) CHK(tdda_model_set(m, TDDA_VEC_AL, j, al));
        }
    }
    (void)n3;
    CHK(model_apply(m, precision, al ? 1 : 0));
    return tdda_model_get(m, TDDA_VEC_Y, -1, y);
}

void tdda_solve_opts_default(tdda_solve_opts* o)
{
    if (!o) return;
    o->max_it = 100000;
    o->tol = 1e-8;
    o->mode = TDDA_MODE_FP64;
    o->x0_zero = 0;
    o->inner_tol = 1e-4;
    o->inner_max_it = 2000;
    o->max_outer = 8;
    o->check_every = 0;
    o->true_residual = 0;
    o->precond = TDDA_PRECOND_NONE;
}

int tdda_model_solve(tdda_model m, const tdda_solve_opts* opts, tdda_solve_info* info)
{
    if (!m || !opts || !info) return fail(TDDA_ERR_ARG, "null argument");
    HIPCHK(hipSetDevice(m->op->ctx->device));
    if (!m->v64.x) return fail(TDDA_ERR_STATE, "solves need an operator built with TDDA_C128");
    // pinned status mirrors and events on first use (models that only the
    // request loop drives never need them)
    if (!m->h_st && hipHostMalloc((void**)&m->h_st, sizeof(SolverState) * 3 * m->nrhs, hipHostMallocDefault) != hipSuccess) {
        m->h_st = nullptr;
        return fail(TDDA_ERR_OOM, "pinned host state");
    }
    if (!m->h_poll &&
        hipHostMalloc((void**)&m->h_poll, sizeof(SolverState) * 2 * m->nrhs, hipHostMallocDefault) != hipSuccess) {
        m->h_poll = nullptr;
        return fail(TDDA_ERR_OOM, "pinned host state");
    }
    for (int k = 0; k < 2; k++)
        if (!m->ev[k] && hipEventCreateWithFlags(&m->ev[k], hipEventDisableTiming) != hipSuccess) {
            m->ev[k] = nullptr;
            return fail(TDDA_ERR_HIP, "event creation");
        }
    SolveSet S{m->op->ctx->stream, m->nrhs, &m->j64, &m->jo, m->v32.x ? &m->j32 : nullptr, m->h_st, m->h_poll, m->ev};
    return solve_set(S, opts, info);
}

// ---------------------------------------------------------------------------
// Batches: the right-hand sides of several models (possibly different
// structures, grids and wavelengths) solved together -- every pass is one
// launch over the tiles of all of them.
// ---------------------------------------------------------------------------
int tdda_batch_create(tdda_model* models, int nmodels, tdda_batch* out)
{
    if (!models || !out || nmodels <= 0) return fail(TDDA_ERR_ARG, "bad model list");
    tdda_ctx ctx = models[0] ? models[0]->op->ctx : nullptr;
    bool all32 = true;
    int nj = 0;
    for (int i = 0; i < nmodels; i++) {
        if (!models[i]) return fail(TDDA_ERR_ARG, "null model");
        if (models[i]->op->ctx != ctx) return fail(TDDA_ERR_ARG, "models of one batch must share a context");
        if (!models[i]->v64.x) return fail(TDDA_ERR_STATE, "batched models need complex128 storage");
        all32 = all32 && models[i]->v32.x != nullptr;
        nj += models[i]->nrhs;
    }
    HIPCHK(hipSetDevice(ctx->device));
    tdda_batch b = new tdda_batch_s();
    b->ctx = ctx;
    ctx->refs++;
    b->nj = nj;
    for (int i = 0; i < nmodels; i++) {
        models[i]->refs++;
        b->models.push_back(models[i]);
        for (int j = 0; j < models[i]->nrhs; j++) {
            b->j64.h.push_back(models[i]->j64.h[j]);
            b->jo.h.push_back(models[i]->jo.h[j]);
            if (all32) b->j32.h.push_back(models[i]->j32.h[j]);
        }
    }
    b->has32 = all32;
    auto bail = [&](int rc) {
        tdda_batch_destroy(b);
        return rc;
    };
    int rc;
    if ((rc = finalize_table<double>(b->allocs, ctx, b->j64)) != TDDA_OK) return bail(rc);
    if ((rc = finalize_table<double>(b->allocs, ctx, b->jo)) != TDDA_OK) return bail(rc);
    if (all32 && (rc = finalize_table<float>(b->allocs, ctx, b->j32)) != TDDA_OK) return bail(rc);
    if (hipHostMalloc((void**)&b->h_st, sizeof(SolverState) * 3 * nj, hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void**)&b->h_poll, sizeof(SolverState) * 2 * nj, hipHostMallocDefault) != hipSuccess)
        return bail(fail(TDDA_ERR_OOM, "pinned host state"));
    for (int k = 0; k < 2; k++)
        if (hipEventCreateWithFlags(&b->ev[k], hipEventDisableTiming) != hipSuccess)
            return bail(fail(TDDA_ERR_HIP, "event creation"));
    if (hipStreamSynchronize(ctx->stream) != hipSuccess) return bail(fail(TDDA_ERR_HIP, "batch set-up"));
    *out = b;
    return TDDA_OK;
}

int tdda_batch_destroy(tdda_batch b)
{
    if (!b) return TDDA_OK;
    (void)hipSetDevice(b->ctx->device);
    (void)hipStreamSynchronize(b->ctx->stream);
    for (void* p : b->allocs) ctx_free(b->ctx, p);
    if (b->h_st) (void)hipHostFree(b->h_st);
    if (b->h_poll) (void)hipHostFree(b->h_poll);
    for (int k = 0; k < 2; k++)
        if (b->ev[k]) (void)hipEventDestroy(b->ev[k]);
    for (tdda_model m : b->models) model_release(m);
    tdda_ctx ctx = b->ctx;
    delete b;
    ctx_release(ctx);
    return TDDA_OK;
}

int tdda_batch_size(tdda_batch b) { return b ? b->nj : 0; }

int tdda_batch_solve(tdda_batch b, const tdda_solve_opts* opts, tdda_solve_info* info)
{
    if (!b || !opts || !info) return fail(TDDA_ERR_ARG, "null argument");
    HIPCHK(hipSetDevice(b->ctx->device));
    SolveSet S{b->ctx->stream, b->nj, &b->j64, &b->jo, b->has32 ? &b->j32 : nullptr, b->h_st, b->h_poll, b->ev};
    return solve_set(S, opts, info);
}

int tdda_batch_time_passes(tdda_batch b, int precision, int reps, double* ms5)
{
    if (!b || !ms5 || reps <= 0) return fail(TDDA_ERR_ARG, "bad argument");
    HIPCHK(hipSetDevice(b->ctx->device));
    hipStream_t s = b->ctx->stream;
    if (precision == TDDA_C128) return time_passes_t<double>(s, b->j64, reps, ms5);
    if (precision == TDDA_C64) {
        if (!b->has32) return fail(TDDA_ERR_STATE, "batch has no complex64 storage");
        for (int j = 0; j < b->nj; j++) {
            const size_t n3 = 3 * (size_t)b->j64.h[j].N;
            k_demote<<<(unsigned)((n3 + kThreads - 1) / kThreads), kThreads, 0, s>>>(b->j64.h[j].x, b->j32.h[j].x, n3);
        }
        return time_passes_t<float>(s, b->j32, reps, ms5);
    }
    return fail(TDDA_ERR_ARG, "precision must be TDDA_C128 or TDDA_C64");
}

int tdda_batch_time_pass(tdda_batch b, int precision, int pass, int reps, double* ms)
{
    if (!b || !ms || reps <= 0 || pass < 0 || pass > 4) return fail(TDDA_ERR_ARG, "bad argument");
    HIPCHK(hipSetDevice(b->ctx->device));
    hipStream_t s = b->ctx->stream;
    if (precision == TDDA_C128) return time_pass_t<double>(s, b->j64, pass, reps, ms);
    if (precision == TDDA_C64) {
        if (!b->has32) return fail(TDDA_ERR_STATE, "batch has no complex64 storage");
        for (int j = 0; j < b->nj; j++) {
            const size_t n3 = 3 * (size_t)b->j64.h[j].N;
            k_demote<<<(unsigned)((n3 + kThreads - 1) / kThreads), kThreads, 0, s>>>(b->j64.h[j].x, b->j32.h[j].x, n3);
        }
        return time_pass_t<float>(s, b->j32, pass, reps, ms);
    }
    return fail(TDDA_ERR_ARG, "precision must be TDDA_C128 or TDDA_C64");
}

int tdda_batch_apply(tdda_batch b, int precision)
{
    if (!b) return fail(TDDA_ERR_ARG, "null batch");
    HIPCHK(hipSetDevice(b->ctx->device));
    hipStream_t s = b->ctx->stream;
    if (precision == TDDA_C128) {
        launch_apply<double>(b->j64, b->nj, IN_PLAIN, OUT_PLAIN, 1, 0, s);
    } else if (precision == TDDA_C64) {
        if (!b->has32) return fail(TDDA_ERR_STATE, "batch has no complex64 storage");
        for (int j = 0; j < b->nj; j++) {
            const size_t n3 = 3 * (size_t)b->j64.h[j].N;
            const unsigned blocks = (unsigned)((n3 + kThreads - 1) / kThreads);
            k_demote<<<blocks, kThreads, 0, s>>>(b->j64.h[j].x, b->j32.h[j].x, n3);
        }
        launch_apply<float>(b->j32, b->nj, IN_PLAIN, OUT_PLAIN, 1, 0, s);
        for (int j = 0; j < b->nj; j++) {
            const size_t n3 = 3 * (size_t)b->j64.h[j].N;
            const unsigned blocks = (unsigned)((n3 + kThreads - 1) / kThreads);
            HIPCHK(hipMemsetAsync(b->j64.h[j].ap, 0, n3 * sizeof(Cx<double>), s));
            k_accum<<<blocks, kThreads, 0, s>>>(b->j64.h[j].ap, b->j32.h[j].ap, n3);
        }
    } else {
        return fail(TDDA_ERR_ARG, "precision must be TDDA_C128 or TDDA_C64");
    }
    CHK(check_launch());
    HIPCHK(hipStreamSynchronize(s));
    return TDDA_OK;
}

}  // extern "C"

#include "tdda_serve.inc"
#include "tdda_design.inc"
#include "tdda_objective.inc"
#include "tdda_tune.inc"
#include "tdda_noalloc.inc"
