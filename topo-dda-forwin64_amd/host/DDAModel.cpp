#include "DDAModel.h"

#include <chrono>
#include <iostream>
#include <string>

namespace topodda {

using namespace std::chrono;

DDAModel::DDAModel(AProductCore* Core_, Vector3d n_K_, double E0_, Vector3d n_E0_, bool verbose_)
    : Core(Core_), E0(E0_), n_E0(n_E0_), n_K(n_K_), verbose(verbose_)
{
    const int N = Core->get_N();
    P = VectorXcd::Zero(3 * (std::size_t)N);
    P_max = P;
    Einternal = VectorXcd::Zero(3 * (std::size_t)N);
    compute_E();
    al = VectorXcd::Zero(3 * (std::size_t)N);
    diel = VectorXcd::Zero(3 * (std::size_t)N);
    UpdateAlpha();
    al_max = al;
    RResult = *Core->get_R();  // DDAModel.cpp:39-40
    check(tdda_model_create(Core->handle(), 1, &dev));
}

DDAModel::DDAModel(AProductCore* Core_, Vector3d n_K_, double E0_, Vector3d n_E0_, VectorXi* RResult_, bool verbose_)
    : DDAModel(Core_, n_K_, E0_, n_E0_, verbose_)
{
    if (!RResult_) throw std::invalid_argument("RResult is null");
    if (RResult_->size() % 3 != 0) throw std::invalid_argument("RResult does not have a size with 3*integer");
    RResultSwitch = true;
    RResult = *RResult_;
    EResult = VectorXcd::Zero(RResult.size());
}

DDAModel::~DDAModel()
{
    if (dev) tdda_model_destroy(dev);
}

void DDAModel::compute_E()
{
    const int N = Core->get_N();
    E = VectorXcd::Zero(3 * (std::size_t)N);
    check(tdda_incident_field(N, Core->get_R()->data(), Core->get_K(), Core->get_d(), E0, n_K.v, n_E0.v,
                              reinterpret_cast<double*>(E.data())));
}

void DDAModel::UpdateAlpha()
{
    VectorXd* diel_old = Core->get_diel_old();
    VectorXcd* mat = Core->get_material();
    check(tdda_polarizability((int)diel_old->size(), diel_old->data(), (int)mat->size(),
                              reinterpret_cast<const double*>(mat->data()), Core->get_lam(), Core->get_K(),
                              Core->get_d(), n_E0.v, n_K.v, reinterpret_cast<double*>(diel.data()),
                              reinterpret_cast<double*>(al.data())));
}

void DDAModel::UpdateAlphaSingle(int idx)
{
    // the reference recomputes component 3*idx and copies it to all three
    VectorXd* diel_old = Core->get_diel_old();
    VectorXcd* mat = Core->get_material();
    double dv[2], av[2];
    check(tdda_polarizability(1, diel_old->data() + 3 * (std::size_t)idx, (int)mat->size(),
                              reinterpret_cast<const double*>(mat->data()), Core->get_lam(), Core->get_K(),
                              Core->get_d(), n_E0.v, n_K.v, dv, av));
    for (int c = 0; c < 3; c++) {
        diel(3 * (std::size_t)idx + c) = cd(dv[0], dv[1]);
        al(3 * (std::size_t)idx + c) = cd(av[0], av[1]);
    }
}

void DDAModel::change_E(VectorXcd E_) { E = E_; }

void DDAModel::E_from_device()
{
    check(tdda_model_get(dev, TDDA_VEC_E, 0, reinterpret_cast<double*>(E.data())));
    E_dev = E;
}

void DDAModel::reset_E() { compute_E(); }

void DDAModel::InitializeP(VectorXcd& Initializer) { P = Initializer; }

void DDAModel::sync_to_device()
{
    if (al_dev != al) {
        check(tdda_model_set(dev, TDDA_VEC_AL, 0, reinterpret_cast<const double*>(al.data())));
        al_dev = al;
    }
    if (E_dev != E) {
        check(tdda_model_set(dev, TDDA_VEC_E, 0, reinterpret_cast<const double*>(E.data())));
        E_dev = E;
    }
    check(tdda_model_set(dev, TDDA_VEC_P, 0, reinterpret_cast<const double*>(P.data())));
}

void DDAModel::bicgstab(int MAX_ITERATION, double MAX_ERROR)
{
    if (verbose) std::cout << "--------------Calculation start. Iterative method used: BICGSTAB---------------\n\n";
    const auto t_start = high_resolution_clock::now();
    sync_to_device();
    tdda_solve_opts o;
    tdda_solve_opts_default(&o);
    o.max_it = MAX_ITERATION;
    o.tol = MAX_ERROR;
    o.mode = solver_mode;
    o.precond = solver_precond;
    o.x0_zero = 0;
    check(tdda_model_solve(dev, &o, &info));
    check(tdda_model_get(dev, TDDA_VEC_P, 0, reinterpret_cast<double*>(P.data())));
    const auto t_end = high_resolution_clock::now();
    time = (int)duration_cast<milliseconds>(t_end - t_start).count();
    if (info.status == TDDA_SOLVE_CONVERGED || info.status == TDDA_SOLVE_BREAKDOWN) {
        // ITERATION and Error are only set on these two exits (DDAModel.cpp:193-216, 232-254)
        ITERATION = info.iterations;
        Error = info.error;
        if (verbose) {
            if (info.status == TDDA_SOLVE_BREAKDOWN)
                std::cout << "----------------------------r0dotrl==(0,0), nan is going to occur so stop now"
                             "------------------------\n";
            std::cout << "--------------Calculation finished. Duration: " << time / 1000.0
                      << "s.-------------------\n";
            std::cout << "              Error: " << Error << "\n              Iteration: " << ITERATION << "\n\n";
        }
    } else {
        std::cout << "                ERROR:does not converge in " << MAX_ITERATION << " iterations" << std::endl;
    }
}

void DDAModel::update_E_in_structure()
{
    check(tdda_local_field((int)P.size(), reinterpret_cast<const double*>(al.data()),
                           reinterpret_cast<const double*>(P.data()), reinterpret_cast<const double*>(diel.data()),
                           reinterpret_cast<double*>(Einternal.data())));
}

void DDAModel::solve_E()
{
    if (!RResultSwitch) {
        EResult = Einternal;
        return;
    }
    check(tdda_model_set(dev, TDDA_VEC_P, 0, reinterpret_cast<const double*>(P.data())));
    check(tdda_field_at_points(dev, 0, RResult.data(), (int)(RResult.size() / 3), n_K.v, n_E0.v, E0,
                               reinterpret_cast<double*>(EResult.data())));
}

void DDAModel::output_to_file()
{
    Core->get_CStr()->output_to_file();
    check(tdda_write_model_results("Model_results.txt", TDDA_RESULTS_FULL, reinterpret_cast<const double*>(P.data()),
                                   (int)P.size(), reinterpret_cast<const double*>(EResult.data()),
                                   (int)EResult.size(), RResult.data(), (int)RResult.size(), n_K.v, n_E0.v));
}

void DDAModel::output_to_file(std::string save_position, int iteration, int ModelLabel)
{
    const std::string name = save_position + "Model_results" + std::to_string(ModelLabel) + "it" +
                             std::to_string(iteration) + ".txt";
    check(tdda_write_model_results(name.c_str(), TDDA_RESULTS_E_P, reinterpret_cast<const double*>(P.data()),
                                   (int)P.size(), reinterpret_cast<const double*>(EResult.data()),
                                   (int)EResult.size(), nullptr, 0, nullptr, nullptr));
}

void DDAModel::output_to_file(std::string save_position, int iteration)
{
    const std::string name = save_position + "Model_results" + "it" + std::to_string(iteration) + ".txt";
    check(tdda_write_model_results(name.c_str(), TDDA_RESULTS_E, nullptr, 0,
                                   reinterpret_cast<const double*>(EResult.data()), (int)EResult.size(), nullptr, 0,
                                   nullptr, nullptr));
}

VectorXcd DDAModel::Aproductwithalb(VectorXcd& b)
{
    if (b.size() != al.size()) std::cout << "In Aproductwithalb, bsize not equal to alsize" << std::endl;
    VectorXcd y(b.size());
    check(tdda_aproduct(Core->handle(), reinterpret_cast<const double*>(b.data()), reinterpret_cast<double*>(y.data()),
                        1, reinterpret_cast<const double*>(al.data()), 0, TDDA_C128));
    return y;
}

}  // namespace topodda
