// EvoDDAModel.h -- the solve call surface of the reference's adjoint
// optimiser (EvoDDAModel.h:10-71): the forward solve of every model and the
// adjoint solve with the same operator, in the reference's order, plus the
// gradient assembly that consumes them.  Objectives, finite-difference
// derivatives and the Adam/Adagrad update are host-side design code and stay
// with the caller.
#pragma once
#include <vector>

#include "DDAModel.h"

namespace topodda {

class EvoDDAModel {
public:
    EvoDDAModel(CoreStructure* CStr, std::vector<DDAModel*> allModel, bool HaveOriginHeritage = true,
                bool HaveAdjointHeritage = true);

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

private:
    CoreStructure* CStr;
    std::vector<DDAModel*> allModel;
    int ModelNum;
    bool HaveOriginHeritage, HaveAdjointHeritage;
    std::vector<VectorXcd> PforOrigin, PforAdjoint;
    int TotalOriginIt = 0, TotalAdjointIt = 0;
};

}  // namespace topodda
