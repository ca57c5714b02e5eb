// tdda_launch.hpp -- host launchers of the operator passes.  Each launcher
// picks the kernel specialised for the pass's FFT length when one is
// compiled (lengths 2^a, 3 * 2^a, 5 * 2^a from 8 to 1024; z passes up to
// 128) and the runtime-plan kernel otherwise.  Definitions live in
// tdda_pass_f64.hip / tdda_pass_f32.hip so both precisions build in parallel.
#pragma once
#include <hip/hip_runtime.h>

#include "tdda_kernels.hpp"

namespace tdda {

struct PassLaunch {
    dim3 grid;  // tiles: the launcher caps it at the resident workgroups
    size_t lds;
    hipStream_t stream;
    int ncu;
    int threads = kThreads;  // workgroup size of the pass kernel
};

template <typename T>
void launch_p1(int L, int mode, const PassLaunch& pl, const JobDesc<T>* jobs, const int* tstart, int nj);
template <typename T>
void launch_p5(int L, int mode, int diag, const PassLaunch& pl, const JobDesc<T>* jobs, const int* tstart, int nj);
template <typename T>
void launch_py(int L, bool inv, int solver, const PassLaunch& pl, const JobDesc<T>* jobs, const int* tstart, int nj,
               const int4* tmap = nullptr);
template <typename T>
void launch_px(int L, int solver, const PassLaunch& pl, const JobDesc<T>* jobs, const int* tstart, int nj,
               const int4* tmap = nullptr);
// Fused z + y passes (tdda_zy.hpp) of an operator whose y length has a
// register split up to 128 points (zy_ylen_ok) and whose z length is one of
// the register DFT lengths (zy_zlen_ok): k_pzy replaces P1 and P2, k_pyz P4
// and P5.
template <typename T>
void launch_pzy(int L, int mode, const PassLaunch& pl, const JobDesc<T>* jobs, const int* tstart, int nj,
                const int2* tmap = nullptr);
template <typename T>
void launch_pyz(int L, int mode, int diag, const PassLaunch& pl, const JobDesc<T>* jobs, const int* tstart, int nj,
                const int2* tmap = nullptr);
// Launch-group key bits of the register-resident x and y passes: launch_px
// runs k_pxr<L> for L | kPxrKey (with 512 threads per workgroup when
// kPxrWide is also set), launch_py k_pyr<L>.
constexpr int kPxrKey = 1 << 20;
constexpr int kPxrWide = 1 << 21;
// true when a specialised kernel exists for this length and pass kind
// (0: z pass, 1: y pass, 2: x pass, 3: register-resident x pass)
template <typename T>
bool pass_length_compiled(int kind, int L);

}  // namespace tdda
