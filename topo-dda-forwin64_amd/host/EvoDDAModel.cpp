#include "EvoDDAModel.h"

#include <iostream>
#include <list>
#include <stdexcept>

namespace topodda {

EvoDDAModel::EvoDDAModel(CoreStructure* CStr_, std::vector<DDAModel*> allModel_, bool origin_heritage,
                         bool adjoint_heritage)
    : CStr(CStr_), allModel(allModel_), ModelNum((int)allModel_.size()), HaveOriginHeritage(origin_heritage),
      HaveAdjointHeritage(adjoint_heritage)
{
    for (int i = 0; i < ModelNum; i++)
        if (allModel[i]->get_Core()->get_CStr() != CStr)
            std::cout << "The DDAModel number: " << i << " does not share the same CoreStructure" << std::endl;
    const std::size_t n3 = 3 * (std::size_t)CStr->get_N();
    for (int i = 0; i < ModelNum; i++) {
        PforOrigin.push_back(VectorXcd::Zero(n3));
        PforAdjoint.push_back(VectorXcd::Zero(n3));
    }
}

void EvoDDAModel::ForwardSolve(int MAX_ITERATION, double MAX_ERROR)
{
    for (int i = 0; i < ModelNum; i++) {
        allModel[i]->InitializeP(PforOrigin[i]);
        allModel[i]->bicgstab(MAX_ITERATION, MAX_ERROR);
        if (HaveOriginHeritage) PforOrigin[i] = *(allModel[i]->get_P());
        allModel[i]->update_E_in_structure();
        TotalOriginIt += allModel[i]->get_ITERATION();
    }
}

VectorXcd EvoDDAModel::AdjointSolve(int i, const VectorXcd& devp, int MAX_ITERATION, double MAX_ERROR)
{
    if (i < 0 || i >= ModelNum) throw std::out_of_range("model index");
    allModel[i]->change_E(devp);
    allModel[i]->InitializeP(PforAdjoint[i]);
    allModel[i]->bicgstab(MAX_ITERATION, MAX_ERROR);
    if (HaveAdjointHeritage) PforAdjoint[i] = *(allModel[i]->get_P());
    VectorXcd lambdaT = *(allModel[i]->get_P());
    allModel[i]->reset_E();
    TotalAdjointIt += allModel[i]->get_ITERATION();
    return lambdaT;
}

VectorXd EvoDDAModel::Gradient(const VectorXd& devx, const VectorXcd& lambdaT, const VectorXcd& Adevxp,
                               const VectorXi& geometryPara, const VectorXi& Free, int n_para_all)
{
    const int n_para = (int)Free.size();
    const int N = (int)geometryPara.size();
    std::vector<std::list<int>> Paratogeometry(n_para_all);
    for (int i = 0; i < N; i++) Paratogeometry[geometryPara(i)].push_back(i);
    VectorXd g = VectorXd::Zero(n_para);
    for (int i = 0; i < n_para; i++) {
        cd acc(0.0, 0.0);
        for (int position : Paratogeometry[Free(i)])
            for (int c = 0; c < 3; c++) acc += lambdaT(3 * (std::size_t)position + c) * Adevxp(3 * (std::size_t)position + c);
        g(i) = devx(i) - acc.real();
    }
    return g;
}

}  // namespace topodda
