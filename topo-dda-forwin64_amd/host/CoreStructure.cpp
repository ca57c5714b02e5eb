#include "CoreStructure.h"

#include "AProductCore.h"

#include <tuple>

namespace topodda {

CoreStructure::CoreStructure(int Nx_, int Ny_, int Nz_, double d_, const VectorXi& R_, const VectorXd& diel_old_)
    : N((int)(R_.size() / 3)), Nx(Nx_), Ny(Ny_), Nz(Nz_), d(d_), R(R_), diel_old(diel_old_), diel_old_max(diel_old_)
{
    if (R_.size() % 3 != 0) throw std::invalid_argument("R must hold 3N coordinates");
    if (diel_old_.size() != R_.size()) throw std::invalid_argument("diel_old must hold 3N values");
}

CoreStructure::CoreStructure(SpacePara* spacepara_, double d_) : spacepara(spacepara_), d(d_)
{
    if (!spacepara) throw std::invalid_argument("CoreStructure: null SpacePara");
    std::tie(Nx, Ny, Nz, N) = spacepara->get_space()->get_Ns();
    R = spacepara->get_geometry();
    diel_old = VectorXd::Zero(3 * (std::size_t)N);
    diel_from_para();
    diel_old_max = VectorXd::Zero(3 * (std::size_t)N);
}

void CoreStructure::diel_from_para()
{
    const VectorXi* gp = spacepara->get_geometryPara();
    const VectorXd* Para = spacepara->get_Para();
    for (int i = 0; i < N; i++) {
        const double v = (*Para)((*gp)(i));
        for (int c = 0; c < 3; c++) diel_old(3 * (std::size_t)i + c) = v;
    }
}

void CoreStructure::UpdateStr(VectorXd step, int current_it, int Max_it)
{
    if (!spacepara) throw std::logic_error("CoreStructure::UpdateStr(step, ...) needs the SpacePara constructor");
    VectorXd& Para = *spacepara->get_Para();
    const VectorXi& Free = *spacepara->get_Free();
    const std::size_t n = Free.size();
    if (step.size() != n) throw std::invalid_argument("CoreStructure::UpdateStr: step.size != FreePara.size");
    auto clamp01 = [](double& v) {
        if (v >= 1) v = 1;
        if (v <= 0) v = 0;
    };
    if (spacepara->get_Filter()) {
        VectorXd& origin = *spacepara->get_Para_origin();
        VectorXd& filtered = *spacepara->get_Para_filtered();
        FilterOption* fo = spacepara->get_Filterstats();
        fo->update_beta(current_it, Max_it);
        for (std::size_t i = 0; i < n; i++) {
            origin(Free(i)) += step(i);
            clamp01(origin(Free(i)));
        }
        const std::vector<std::vector<WeightPara>>& W = *spacepara->get_FreeWeight();
        for (std::size_t i = 0; i < n; i++) {
            double num = 0.0, den = 0.0;
            for (const WeightPara& w : W[i]) {
                num += w.weight * origin(w.position);
                den += w.weight;
            }
            filtered(Free(i)) = num / den;
            Para(Free(i)) = fo->SmoothDensity(filtered(Free(i)));
        }
    } else {
        for (std::size_t i = 0; i < n; i++) {
            Para(Free(i)) += step(i);
            clamp01(Para(Free(i)));
        }
    }
    diel_from_para();
}

VectorXi CoreStructure::build_a_bulk(int Nx, int Ny, int Nz)
{
    VectorXi r = VectorXi::Zero(3 * (std::size_t)Nx * Ny * Nz);
    for (int x = 0; x < Nx; x++)
        for (int y = 0; y < Ny; y++)
            for (int z = 0; z < Nz; z++) {
                const std::size_t pos = (std::size_t)z + Nz * ((std::size_t)y + Ny * x);
                r(3 * pos) = x;
                r(3 * pos + 1) = y;
                r(3 * pos + 2) = z;
            }
    return r;
}

CoreStructure CoreStructure::Bulk(int Nx, int Ny, int Nz, double d, const VectorXd& diel_old)
{
    return CoreStructure(Nx, Ny, Nz, d, build_a_bulk(Nx, Ny, Nz), diel_old);
}

void CoreStructure::UpdateStr(const VectorXd& diel_old_new)
{
    if (diel_old_new.size() != diel_old.size()) throw std::invalid_argument("diel_old size mismatch");
    diel_old = diel_old_new;
}

void CoreStructure::UpdateStrSingle(int idx, double value)
{
    if (idx < 0 || idx >= N) throw std::out_of_range("dipole index");
    diel_old(3 * idx) = value;
    diel_old(3 * idx + 1) = value;
    diel_old(3 * idx + 2) = value;
}

void CoreStructure::output_to_file()
{
    check(tdda_write_structure("CoreStructure.txt", Nx, Ny, Nz, N, R.data(), diel_old.data(), d, 1));
}

void CoreStructure::output_to_file(std::string save_position, int iteration, std::string mode)
{
    const std::string name = save_position + "CoreStructure" + std::to_string(iteration) + ".txt";
    check(tdda_write_structure(name.c_str(), Nx, Ny, Nz, N, R.data(), diel_old.data(), d, mode == "normal" ? 1 : 0));
}

}  // namespace topodda
