#include "ObjDDAModel.h"

#include <cmath>
#include <stdexcept>

namespace topodda {

namespace {

double smooth_density(double input, double ita, double beta)  // ObjDDAModel.cpp:5-17
{
    if (input <= ita && input >= 0.0) return ita * (std::exp(-beta * (1 - input / ita)) - (1 - input / ita) * std::exp(-beta));
    if (input > ita && input <= 1.0)
        return (1 - ita) * (1 - std::exp(-beta * (input - ita) / (1 - ita)) + (input - ita) / (1 - ita) * std::exp(-beta)) +
               ita;
    throw std::out_of_range("SmoothDensity: input out of range");
}

}  // namespace

ObjPointEDDAModel::ObjPointEDDAModel(std::vector<double> parameters, DDAModel* model_) : model(model_)
{
    if (parameters.size() < 3) throw std::invalid_argument("PointE needs 3 parameters (x, y, z)");
    for (int a = 0; a < 3; a++) r0[a] = parameters[a];
    Have_Devx = false;
    const Vector3d nK = model->get_nK(), nE = model->get_nE0();
    const double K = model->get_K(), E0 = model->get_E0();
    const double ph = K * (nK(0) * r0[0] + nK(1) * r0[1] + nK(2) * r0[2]);
    for (int a = 0; a < 3; a++) E_ext[a] = E0 * nE(a) * cd(std::cos(ph), std::sin(ph));
    Reset();
}

double ObjPointEDDAModel::GetVal()
{
    model->to_device();
    const Vector3d nK = model->get_nK(), nE = model->get_nE0();
    check(tdda_objective_point_e(model->device_model(), 0, r0, nK.v, nE.v, model->get_E0(), &value, nullptr, 0));
    host_valid = false;
    return value;
}

void ObjPointEDDAModel::Reset()
{
    for (int a = 0; a < 3; a++) E_sum[a] = E_ext[a];
}

void ObjPointEDDAModel::host_sum()
{
    Reset();
    for (int idx = 0; idx < model->get_N(); idx++) {
        host_valid = true;
        SingleResponse(idx, false);
    }
    host_valid = true;
}

// ObjDDAModel.cpp:50-72
void ObjPointEDDAModel::SingleResponse(int idx, bool deduction)
{
    if (!host_valid) host_sum();
    const VectorXi& R = *model->get_R();
    const VectorXcd& P = *model->get_P();
    const double d = model->get_d();
    const Matrix3cd A = model->get_Core()->A_dic_generator(r0[0] - d * R(3 * (std::size_t)idx),
                                                          r0[1] - d * R(3 * (std::size_t)idx + 1),
                                                          r0[2] - d * R(3 * (std::size_t)idx + 2));
    for (int a = 0; a < 3; a++) {
        const cd t = A(a, 0) * P(3 * (std::size_t)idx) + A(a, 1) * P(3 * (std::size_t)idx + 1) +
                     A(a, 2) * P(3 * (std::size_t)idx + 2);
        E_sum[a] += deduction ? t : -t;
    }
}

double ObjPointEDDAModel::GroupResponse()
{
    if (!host_valid) return value;
    return std::sqrt(std::norm(E_sum[0]) + std::norm(E_sum[1]) + std::norm(E_sum[2]));
}

double ObjPointEDDAModel::AdjointSource()
{
    model->to_device();
    const Vector3d nK = model->get_nK(), nE = model->get_nE0();
    tdda_model m = model->device_model();
    check(tdda_objective_point_e(m, 0, r0, nK.v, nE.v, model->get_E0(), &value, m, 0));
    model->E_from_device();
    host_valid = false;
    return value;
}

ObjIntegratedEDDAModel::ObjIntegratedEDDAModel(std::vector<double> p, DDAModel* model_) : model(model_)
{
    if (p.size() < 9) throw std::invalid_argument("IntegratedE needs 9 parameters");
    powNum = (int)std::round(p[0]);
    for (int k = 0; k < 6; k++) box[k] = (int)std::round(p[1 + k]);
    ita = p[7];
    beta = p[8];
    Have_Devx = true;
}

double ObjIntegratedEDDAModel::GetVal()
{
    model->to_device();
    double v = 0.0;
    check(tdda_objective_integrated_e(model->device_model(), 0, powNum, box, ita, beta,
                                      model->get_diel_old()->data(), &v, nullptr, 0));
    // the host sum the incremental updates start from
    Reset();
    for (int idx = 0; idx < model->get_N(); idx++) SingleResponse(idx, false);
    return v;
}

void ObjIntegratedEDDAModel::Reset() { E_int = 0.0; }

// ObjDDAModel.cpp:136-152
void ObjIntegratedEDDAModel::SingleResponse(int idx, bool deduction)
{
    const VectorXi& R = *model->get_R();
    const std::size_t i = 3 * (std::size_t)idx;
    if (((box[0] <= R(i)) <= box[1]) && ((box[2] <= R(i + 1)) <= box[3]) && ((box[4] <= R(i + 2)) <= box[5])) {
        const VectorXcd& al = *model->get_al();
        const VectorXcd& P = *model->get_P();
        const double factor = smooth_density((*model->get_diel_old())(i), ita, beta);
        for (int c = 0; c < 3; c++) {
            const double t = factor * std::pow(std::abs(al(i + c) * P(i + c)), powNum);
            E_int += deduction ? -t : t;
        }
    }
}

double ObjIntegratedEDDAModel::GroupResponse() { return std::log(E_int); }

double ObjIntegratedEDDAModel::AdjointSource()
{
    model->to_device();
    double v = 0.0;
    tdda_model m = model->device_model();
    check(tdda_objective_integrated_e(m, 0, powNum, box, ita, beta, model->get_diel_old()->data(), &v, m, 0));
    model->E_from_device();
    return v;
}

}  // namespace topodda
