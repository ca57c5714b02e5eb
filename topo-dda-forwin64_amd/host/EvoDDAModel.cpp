#include "EvoDDAModel.h"

#include <chrono>
#include <cmath>
#include <filesystem>
#include <fstream>
#include <iostream>
#include <list>
#include <stdexcept>

namespace topodda {

EvoDDAModel::EvoDDAModel(std::string objName_, std::vector<double> objPara_, double epsilon_fix_,
                         bool HavePathRecord_, bool HaveOriginHeritage_, bool HaveAdjointHeritage_,
                         std::string save_position_, CoreStructure* CStr_, std::vector<DDAModel*> allModel_)
    : EvoDDAModel(CStr_, allModel_, HaveOriginHeritage_, HaveAdjointHeritage_)
{
    objName = objName_;
    objPara = objPara_;
    save_position = save_position_;
    epsilon_fix = epsilon_fix_;
    epsilon_tmp = epsilon_fix;
    HavePathRecord = HavePathRecord_;
    if (!CStr->get_spacepara()) throw std::invalid_argument("EvoDDAModel needs a CoreStructure built from a SpacePara");
    gradientsquare = VectorXd::Zero(CStr->get_spacepara()->get_Para()->size());
    for (int i = 0; i < ModelNum; i++) allObj.push_back(ObjFactory(objName, objPara, allModel[i]));
}

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
        PforOriginMax.push_back(VectorXcd::Zero(n3));
        PforAdjointMax.push_back(VectorXcd::Zero(n3));
    }
    Originarray = VectorXd::Zero(ModelNum);
    MaxObjarray = VectorXd::Zero(ModelNum);
}

EvoDDAModel::~EvoDDAModel()
{
    for (ObjDDAModel* o : allObj) delete o;
}

ObjDDAModel* EvoDDAModel::ObjFactory(std::string ObjectName, std::vector<double> ObjectParameters, DDAModel* model)
{
    if (ObjectName == "PointE") return new ObjPointEDDAModel(ObjectParameters, model);
    if (ObjectName == "IntegratedE") return new ObjIntegratedEDDAModel(ObjectParameters, model);
    throw std::invalid_argument("EvoDDAModel: unknown objective " + ObjectName);
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

VectorXcd EvoDDAModel::adjoint(int i, int MAX_ITERATION, double MAX_ERROR, int& it)
{
    DDAModel* m = allModel[i];
    allObj[i]->AdjointSource();  // change_E(devp), EvoDDAModel.cpp:596
    m->InitializeP(PforAdjoint[i]);
    m->bicgstab(MAX_ITERATION, MAX_ERROR);
    if (HaveAdjointHeritage) PforAdjoint[i] = *(m->get_P());
    VectorXcd lambdaT = *(m->get_P());
    m->reset_E();
    it = m->get_ITERATION();
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

std::tuple<VectorXd, VectorXcd> EvoDDAModel::devx_and_Adevxp(double epsilon, DDAModel* CurrentModel, ObjDDAModel* Obj,
                                                            double origin)
{
    SpacePara* sp = CStr->get_spacepara();
    const VectorXi& geometryPara = *sp->get_geometryPara();
    const VectorXd& Para = *sp->get_Para();
    const VectorXi& Free = *sp->get_Free();
    const int N = CurrentModel->get_N();
    VectorXcd& al = *CurrentModel->get_al();
    VectorXcd& P = *CurrentModel->get_P();
    VectorXcd Adevxp = VectorXcd::Zero(3 * (std::size_t)N);
    const int n_para = (int)Free.size();
    VectorXd devx = VectorXd::Zero(n_para);
    std::vector<std::list<int>> Paratogeometry(Para.size());
    for (int i = 0; i < N; i++) Paratogeometry[geometryPara(i)].push_back(i);
    // the reference's two branches (with and without filter) are the same
    for (int i = 0; i < n_para; i++) {
        const int pos = Free(i);
        const double t0 = Para(pos);
        const int sign = t0 >= epsilon ? -1 : 1;
        const double t1 = t0 + sign * epsilon;
        for (int position : Paratogeometry[pos]) {
            const cd alphaorigin = al(3 * (std::size_t)position);
            if (Obj->Have_Devx) Obj->SingleResponse(position, true);
            CStr->UpdateStrSingle(position, t1);
            CurrentModel->UpdateAlphaSingle(position);
            if (Obj->Have_Devx) Obj->SingleResponse(position, false);
            const cd change = (al(3 * (std::size_t)position) - alphaorigin) / (double)(sign * epsilon);
            for (int c = 0; c < 3; c++) Adevxp(3 * (std::size_t)position + c) = change;
        }
        devx(i) = (Obj->GroupResponse() - origin) / (sign * epsilon);
        for (int position : Paratogeometry[pos]) {
            if (Obj->Have_Devx) Obj->SingleResponse(position, true);
            CStr->UpdateStrSingle(position, t0);
            CurrentModel->UpdateAlphaSingle(position);
            if (Obj->Have_Devx) Obj->SingleResponse(position, false);
        }
    }
    for (std::size_t i = 0; i < 3 * (std::size_t)N; i++) Adevxp(i) = Adevxp(i) * P(i);
    return std::make_tuple(devx, Adevxp);
}

VectorXd EvoDDAModel::gradients_filtered(VectorXd gradients, int current_it, int Max_it)
{
    SpacePara* sp = CStr->get_spacepara();
    FilterOption* fo = sp->get_Filterstats();
    const VectorXd& Para_filtered = *sp->get_Para_filtered();
    const VectorXi& Free = *sp->get_Free();
    fo->update_beta(current_it, Max_it);
    const double gbeta = fo->get_beta(), gita = fo->get_ita();
    const int NFpara = (int)gradients.size();
    const std::vector<std::vector<WeightPara>>& FW = *sp->get_FreeWeight();
    if ((std::size_t)NFpara != FW.size()) throw std::logic_error("gradients_filtered: NFpara != FreeWeight.size()");
    VectorXd result = VectorXd::Zero(NFpara);
    for (int i = 0; i < NFpara; i++) {
        double sum_weight = 0.0;
        for (const WeightPara& w : FW[i]) sum_weight += w.weight;
        for (const WeightPara& w : FW[i])
            if (w.position < NFpara) result(w.position) += w.weight / sum_weight;
    }
    for (int i = 0; i < NFpara; i++) {
        const int position = Free(i);
        const double ptofd = PtoFderivative(Para_filtered(position), gbeta, gita);
        result(position) = result(position) * ptofd * gradients(position);
    }
    return result;
}

void EvoDDAModel::EvoOptimization(int MAX_ITERATION, double MAX_ERROR, int MAX_ITERATION_EVO, std::string method,
                                  double start_num)
{
    using clk = std::chrono::high_resolution_clock;
    if (allObj.empty()) throw std::logic_error("EvoOptimization needs the objective constructor");
    namespace fs = std::filesystem;
    if (!save_position.empty()) fs::create_directories(save_position);
    fs::create_directories(save_position + "CoreStructure");
    fs::create_directories(save_position + "Model_output");
    std::ofstream convergence(save_position + "convergence.txt");
    std::ofstream Originiterations(save_position + "Originiterations.txt");
    std::ofstream Adjointiterations(save_position + "Adjointiterations.txt");
    // Adam (EvoDDAModel.cpp:408-411)
    const double beta1 = 0.9;
    const double beta2 = 1 - (1 - beta1) * (1 - beta1);
    VectorXd V, S;
    const double epsilon_partial = 0.001;
    SpacePara* spacepara = CStr->get_spacepara();
    for (int iteration = 0; iteration < MAX_ITERATION_EVO; iteration++) {
        VectorXd objarray = VectorXd::Zero(ModelNum);
        auto t0 = clk::now();
        CStr->output_to_file(save_position + "CoreStructure/", (int)(iteration + start_num), "simple");
        output_time += std::chrono::duration_cast<std::chrono::milliseconds>(clk::now() - t0).count();
        for (int i = 0; i < ModelNum; i++) {
            DDAModel* m = allModel[i];
            m->InitializeP(PforOrigin[i]);
            m->bicgstab(MAX_ITERATION, MAX_ERROR);
            if (HaveOriginHeritage) PforOrigin[i] = *m->get_P();
            m->update_E_in_structure();
            m->solve_E();
            t0 = clk::now();
            m->output_to_file(save_position + "Model_output/", (int)(iteration + start_num), i);
            output_time += std::chrono::duration_cast<std::chrono::milliseconds>(clk::now() - t0).count();
            objarray(i) = allObj[i]->GetVal();
            Originiterations << m->get_ITERATION() << std::endl;
            TotalOriginIt += m->get_ITERATION();
        }
        double obj = objarray.sum() / ModelNum;
        history.push_back(obj);
        convergence << obj << " ";
        std::cout << "Obj function at iteration " << iteration << " is " << obj << std::endl;
        VectorXd* diel_old = CStr->get_diel_old();
        VectorXd* diel_old_max = CStr->get_diel_old_max();
        double epsilon = epsilon_fix;
        if (HavePathRecord) {
            // EvoDDAModel.cpp:471-557: back-tracking with step-size control
            if (std::abs(obj - PreviousObj) / PreviousObj <= 0.0001 || epsilon <= 0.0001) CutoffHold += 1;
            else if (CutoffHold > 0) CutoffHold -= 1;
            PreviousObj = obj;
            if (obj < MaxObj) {
                epsilon_tmp = epsilon_tmp / 10;
                Stephold = 0;
                *diel_old = *diel_old_max;
                for (int i = 0; i < ModelNum; i++) {
                    *allModel[i]->get_P() = *allModel[i]->get_P_max();
                    *allModel[i]->get_al() = *allModel[i]->get_al_max();
                    objarray(i) = MaxObjarray(i);
                    PforOrigin[i] = PforOriginMax[i];
                    PforAdjoint[i] = PforAdjointMax[i];
                }
                obj = MaxObj;
            } else {
                if (std::abs(obj - PreviousObj) / PreviousObj <= 0.0001) CutoffHold += 1;
                else if (CutoffHold > 0) CutoffHold -= 1;
                *diel_old_max = *diel_old;
                for (int i = 0; i < ModelNum; i++) {
                    *allModel[i]->get_P_max() = *allModel[i]->get_P();
                    *allModel[i]->get_al_max() = *allModel[i]->get_al();
                    MaxObjarray(i) = objarray(i);
                    PforOriginMax[i] = PforOrigin[i];
                    PforAdjointMax[i] = PforAdjoint[i];
                }
                MaxObj = obj;
                Stephold += 1;
                if (Stephold >= 2) {
                    epsilon_tmp = epsilon_tmp * 10;
                    Stephold = 0;
                }
            }
            epsilon = epsilon_tmp;
        }
        for (int i = 0; i < ModelNum; i++) Originarray(i) = objarray(i);
        convergence << "\n";
        const VectorXi& geometryPara = *spacepara->get_geometryPara();
        const VectorXi& Free = *spacepara->get_Free();
        const int n_para_all = (int)spacepara->get_Para()->size();
        const int n_para = (int)Free.size();
        VectorXd gradients = VectorXd::Zero(n_para);
        for (int i = 0; i < ModelNum; i++) {
            VectorXd devx;
            VectorXcd Adevxp;
            std::tie(devx, Adevxp) = devx_and_Adevxp(epsilon_partial, allModel[i], allObj[i], Originarray(i));
            int it = 0;
            const VectorXcd lambdaT = adjoint(i, MAX_ITERATION, MAX_ERROR, it);
            Adjointiterations << it << std::endl;
            TotalAdjointIt += it;
            gradients = gradients + Gradient(devx, lambdaT, Adevxp, geometryPara, Free, n_para_all);
        }
        gradients = gradients / (double)ModelNum;
        if (spacepara->get_Filter() && iteration >= 1)
            gradients = gradients_filtered(gradients, iteration - 1, MAX_ITERATION_EVO - 1);
        double epsilon_final = epsilon;
        if (method == "Adam" || method == "Adamdecay") {
            const double c1 = 1 - std::pow(beta1, iteration + 1), c2 = 1 - std::pow(beta2, iteration + 1);
            VectorXd g2 = VectorXd::Zero(n_para);
            for (int k = 0; k < n_para; k++) g2(k) = gradients(k) * gradients(k);
            if (iteration == 0) {
                V = gradients * (1 - beta1) / c1;
                S = g2 * (1 - beta2) / c2;
            } else {
                V = V * beta1 + gradients * (1 - beta1) / c1;
                S = S * beta2 + g2 * (1 - beta2) / c2;
            }
            for (int k = 0; k < n_para; k++) gradients(k) = V(k) / (std::sqrt(S(k)) + 0.00000001);
            if (iteration <= 3) epsilon_final = 0.1;
            else if (method == "Adamdecay" && iteration >= 150)
                epsilon_final = epsilon * std::exp(-(iteration - 150) / 10.0);
            else epsilon_final = epsilon;
        }
        if (method == "Adagrad") {
            for (int k = 0; k < n_para; k++) {
                gradientsquare(k) += std::pow(gradients(k), 2) / 100000;
                gradients(k) = gradients(k) / std::sqrt(gradientsquare(k) + 1);
            }
        }
        double mean_abs = 0.0;
        for (int k = 0; k < n_para; k++) mean_abs += std::abs(gradients(k));
        mean_abs = n_para ? mean_abs / n_para : 0.0;
        if (std::abs(mean_abs) < 0.1) gradients = gradients / (std::abs(mean_abs) / 0.1);
        const VectorXd step = gradients * epsilon_final;
        if (spacepara->get_Filter() && spacepara->get_Filterstats()->filterchange(iteration)) spacepara->ChangeFilter();
        CStr->UpdateStr(step, iteration, MAX_ITERATION_EVO - 1);
        for (int i = 0; i < ModelNum; i++) allModel[i]->UpdateAlpha();
    }
    Originiterations << TotalOriginIt << std::endl;
    Adjointiterations << TotalAdjointIt << std::endl;
}

}  // namespace topodda
