// DDAModel.h -- the reference's model class (DDAModel.h:34-74): incident
// field, 1/alpha per component and the BiCGSTAB solve, with the solver state
// resident on the device (libtopodda_hip.so).
//
// Host mirrors of P, E, al and diel are kept coherent: callers may read or
// change *get_P() / *get_al() between solves exactly as EvoDDAModel does
// (EvoDDAModel.cpp:360-389, 489-500); bicgstab() uploads what changed and
// downloads P afterwards.
#pragma once
#include "AProductCore.h"

namespace topodda {

class DDAModel {
public:
    // DDAModel.cpp:12-69
    DDAModel(AProductCore* Core, Vector3d n_K, double E0, Vector3d n_E0, bool verbose = true);
    // DDAModel.cpp:71-129: EResult is evaluated at the lattice points
    // RResult (3 ints each) instead of the sites.
    DDAModel(AProductCore* Core, Vector3d n_K, double E0, Vector3d n_E0, VectorXi* RResult, bool verbose = true);
    ~DDAModel();
    DDAModel(const DDAModel&) = delete;
    DDAModel& operator=(const DDAModel&) = delete;

    // DDAModel.cpp:131-260, run on the device.  P holds x0 on entry.
    void bicgstab(int MAX_ITERATION, double MAX_ERROR);
    void change_E(VectorXcd E_);            // DDAModel.cpp:453-455
    void reset_E();                         // DDAModel.cpp:457-468
    void UpdateAlpha();                     // DDAModel.cpp:470-487
    void UpdateAlphaSingle(int idx);        // DDAModel.cpp:489-528
    void update_E_in_structure();           // DDAModel.cpp:567-575
    // DDAModel.cpp:530-565: EResult = Einternal, or, with RResult, the field
    // at those points by a direct sum over the sites on the device.
    void solve_E();
    // The reference's text result files (DDAModel.cpp:589-697):
    // "Model_results.txt" (P, n_K, n_E0, EResult size, RResult, EResult; also
    // writes "CoreStructure.txt"), save_position + "Model_results" +
    // ModelLabel + "it" + iteration + ".txt" (EResult, P) and save_position +
    // "Model_results" + "it" + iteration + ".txt" (EResult).
    void output_to_file();
    void output_to_file(std::string save_position, int iteration, int ModelLabel);
    void output_to_file(std::string save_position, int iteration);
    VectorXcd Aproductwithalb(VectorXcd& b);  // DDAModel.cpp:577-586
    void InitializeP(VectorXcd& Initializer);  // DDAModel.cpp:731-733

    VectorXcd* get_P() { return &P; }
    Vector3d get_nE0() { return n_E0; }
    Vector3d get_nK() { return n_K; }
    double get_E0() { return E0; }
    VectorXcd* get_Einternal() { return &Einternal; }
    VectorXcd* get_EResult() { return &EResult; }
    AProductCore* get_Core() { return Core; }
    VectorXcd* get_al() { return &al; }
    VectorXcd* get_P_max() { return &P_max; }
    VectorXcd* get_al_max() { return &al_max; }
    VectorXcd* get_E() { return &E; }
    VectorXcd* get_diel() { return &diel; }
    int get_ITERATION() { return ITERATION; }
    double get_Error() { return Error; }
    int get_time() { return time; }
    int get_N() { return Core->get_N(); }
    int get_Nx() { return Core->get_Nx(); }
    int get_Ny() { return Core->get_Ny(); }
    int get_Nz() { return Core->get_Nz(); }
    VectorXi* get_R() { return Core->get_R(); }
    double get_d() { return Core->get_d(); }
    double get_lam() { return Core->get_lam(); }
    double get_K() { return Core->get_K(); }
    VectorXd* get_diel_old() { return Core->get_diel_old(); }
    VectorXcd* get_material() { return Core->get_material(); }

    // Solver selection (not in the reference): TDDA_MODE_FP64 runs the
    // reference algorithm in complex128; TDDA_MODE_MIXED runs complex64 inner
    // solves under FP64 refinement with a true-residual stop test.
    void set_solver_mode(int mode) { solver_mode = mode; }
    // TDDA_PRECOND_JACOBI (not in the reference): BiCGSTAB on A diag(al)^-1,
    // the same residual and solution in far fewer iterations when 1/alpha
    // varies across the sites; TDDA_PRECOND_NONE (default) is the
    // reference's unpreconditioned iteration.
    void set_solver_precond(int precond) { solver_precond = precond; }
    const tdda_solve_info& last_info() const { return info; }
    // The device state (not in the reference): objectives evaluate P there
    // and write adjoint sources into its E.  to_device uploads the host
    // mirrors that changed (al, E) and P; E_from_device refreshes the host
    // mirror of E.
    tdda_model device_model() { return dev; }
    void to_device() { sync_to_device(); }
    void E_from_device();

private:
    void compute_E();
    void sync_to_device();

    AProductCore* Core;
    double E0;
    Vector3d n_E0, n_K;
    VectorXcd P, E, Einternal, EResult, al, diel, P_max, al_max;
    bool RResultSwitch = false;
    VectorXi RResult;
    VectorXcd al_dev, E_dev;  // last values uploaded
    bool verbose;
    int time = 0;
    int ITERATION = 0;
    double Error = 0.0;
    int solver_mode = TDDA_MODE_FP64;
    int solver_precond = TDDA_PRECOND_NONE;
    tdda_solve_info info{};
    tdda_model dev = nullptr;
};

}  // namespace topodda

// The reference's code names these classes without a namespace; define
// TOPODDA_NO_GLOBAL_NAMES to keep them in namespace topodda only.
#ifndef TOPODDA_NO_GLOBAL_NAMES
using topodda::DDAModel;
#endif
