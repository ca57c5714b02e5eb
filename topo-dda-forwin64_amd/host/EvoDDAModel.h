// EvoDDAModel.h -- the reference's adjoint topology optimiser
// (EvoDDAModel.h:10-71, EvoDDAModel.cpp:8-102, 205-391, 393-765,
// 1102-1227).  Per design iteration: the forward solve of every model on the
// device, the objective on the device, the finite-difference design
// derivative of the reference (devx_and_Adevxp, host), the adjoint
// right-hand side in closed form on the device (replacing the reference's
// finite-difference devp), the adjoint solve with the same operator
// (A^T = A), the gradient devx - Re(lambda^T Adevxp) with no conjugation,
// the filter chain rule, Adam / Adamdecay / Adagrad, and
// CoreStructure::UpdateStr + DDAModel::UpdateAlpha.
//
// The lower-level solve surface (ForwardSolve / AdjointSolve / Gradient)
// stays available for callers that run their own loop.
#pragma once
#include <string>
#include <tuple>
#include <vector>

#include "ObjDDAModel.h"

namespace topodda {

class EvoDDAModel {
public:
    // EvoDDAModel.cpp:8-102.  objName "PointE" or "IntegratedE" (objPara as
    // the objective's constructor takes them); the models must share CStr.
    // save_position: directory prefix of the result files ("" for the
    // working directory).
    EvoDDAModel(std::string objName_, std::vector<double> objPara_, double epsilon_fix_, bool HavePathRecord_,
                bool HaveOriginHeritage_, bool HaveAdjointHeritage_, std::string save_position_, CoreStructure* CStr_,
                std::vector<DDAModel*> allModel_);
    // Solve surface only (no objective).
    EvoDDAModel(CoreStructure* CStr, std::vector<DDAModel*> allModel, bool HaveOriginHeritage = true,
                bool HaveAdjointHeritage = true);
    ~EvoDDAModel();
    EvoDDAModel(const EvoDDAModel&) = delete;
    EvoDDAModel& operator=(const EvoDDAModel&) = delete;

    // EvoDDAModel.cpp:393-765 (method "Adam", "Adamdecay" or "Adagrad";
    // any other name takes the plain gradient step).  Writes
    // save_position + "convergence.txt", "Originiterations.txt",
    // "Adjointiterations.txt", and per iteration "CoreStructure/CoreStructure<it>.txt"
    // (design values) and "Model_output/Model_results<i>it<it>.txt" -- the
    // reference writes them under "CoreStructure\" and "Model_output\"; the
    // directories are created here.
    void EvoOptimization(int MAX_ITERATION, double MAX_ERROR, int MAX_ITERATION_EVO, std::string method,
                         double start_num = 0);
    // EvoDDAModel.cpp:205-351: finite-difference derivative of the objective
    // with respect to each free parameter (devx) and dA/dx P on the diagonal
    // (Adevxp), by perturbing each free parameter by epsilon.
    std::tuple<VectorXd, VectorXcd> devx_and_Adevxp(double epsilon, DDAModel* CurrentModel, ObjDDAModel* Obj,
                                                    double origin);
    // EvoDDAModel.cpp:1177-1227: chain rule through the cone filter and
    // SmoothDensity.
    VectorXd gradients_filtered(VectorXd gradients, int current_it, int Max_it);
    ObjDDAModel* ObjFactory(std::string ObjectName, std::vector<double> ObjectParameters, DDAModel* model);
    double get_output_time() const { return output_time; }

    // EvoDDAModel.cpp:433-441: InitializeP(PforOrigin[i]); bicgstab; keep P
    // as the next warm start; update_E_in_structure.
    void ForwardSolve(int MAX_ITERATION, double MAX_ERROR);
    // EvoDDAModel.cpp:596-608: change_E(devp); InitializeP(PforAdjoint[i]);
    // bicgstab (same A, since A^T = A); keep P; lambdaT = P; reset_E.
    VectorXcd AdjointSolve(int i, const VectorXcd& devp, int MAX_ITERATION, double MAX_ERROR);
    // EvoDDAModel.cpp:611-639: devx - Re(sum over the dipoles of each free
    // parameter of lambdaT o Adevxp), no conjugation.
    static VectorXd Gradient(const VectorXd& devx, const VectorXcd& lambdaT, const VectorXcd& Adevxp,
                             const VectorXi& geometryPara, const VectorXi& Free, int n_para_all);

    std::vector<VectorXcd>& get_PforOrigin() { return PforOrigin; }
    std::vector<VectorXcd>& get_PforAdjoint() { return PforAdjoint; }
    int get_TotalOriginIt() const { return TotalOriginIt; }
    int get_TotalAdjointIt() const { return TotalAdjointIt; }
    // Objective values (mean over the models) of the iterations run so far.
    const std::vector<double>& get_objective_history() const { return history; }

private:
    // EvoDDAModel.cpp:586-609 for model i: the adjoint right-hand side on the
    // device, the adjoint solve, lambda^T
    VectorXcd adjoint(int i, int MAX_ITERATION, double MAX_ERROR, int& it);

    double output_time = 0.0;
    CoreStructure* CStr;
    std::vector<DDAModel*> allModel;
    int ModelNum;
    std::string save_position;
    std::vector<VectorXcd> PforOrigin, PforAdjoint, PforOriginMax, PforAdjointMax;
    std::vector<double> objPara;
    std::string objName;
    std::vector<ObjDDAModel*> allObj;
    VectorXd Originarray, MaxObjarray, gradientsquare;
    double MaxObj = 0.0, PreviousObj = 0.0;
    int CutoffHold = 0, Stephold = 0;
    double epsilon_fix = 0.0, epsilon_tmp = 0.0;
    bool HavePathRecord = false, HaveOriginHeritage = true, HaveAdjointHeritage = true;
    int TotalOriginIt = 0, TotalAdjointIt = 0;
    std::vector<double> history;
};

}  // namespace topodda

// The reference's code names these classes without a namespace; define
// TOPODDA_NO_GLOBAL_NAMES to keep them in namespace topodda only.
#ifndef TOPODDA_NO_GLOBAL_NAMES
using topodda::EvoDDAModel;
#endif
